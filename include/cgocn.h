/*
 * cgocn.h -- C ABI of the MI355X ocean tracer engine (libcgocn.so).
 *
 * The engine replaces the tracer half of one cGENIE ocean step, in the
 * coupling order of genie.f90:306-313 (step_goldstein) and 360-381
 * (step_biogem, biogem_tracercoupling):
 *
 *   tstepo            goldstein.f90:2280-2432  -> cgo_tstepo
 *   tstepo_flux, eos  goldstein.f90:2436-2642, 3048-3082
 *   co                goldstein.f90:2657-2777
 *   step_biogem       biogem.f90:528-1877      -> cgo_bio_step: the
 *                     1N1T_PO4MM biology, DOM/particulate remineralisation and
 *                     surface carbonate chemistry of the species carried
 *   biogem_tracercoupling  biogem.f90:1885-2077 -> cgo_tracercoupling
 *   all three in order                         -> cgo_step, cgo_graph_capture
 *
 * and, where a member has them, the steps either side in genie's order:
 *
 *   surflux           embm.f90:2548-3738       -> cgo_surflux (EMBM's surface
 *                     fluxes into ts(:,:,:,maxk+1), goldstein.f90:149-170)
 *   velocity solve    goldstein.f90:198-233    -> before each member's tstepo
 *   air-sea exchange, ATCHEM  biogem_box.f90:27-299, atchem.f90:63-158
 *
 * The reference exposes these as Fortran module subroutines on module state
 * (goldstein_lib.f90:1-180, biogem_lib.f90); there is no C interface
 * upstream.  Host arrays cross this boundary only in the reference's own
 * Fortran layouts (column-major, the lower bounds of its ALLOCATE
 * statements), so a Fortran caller passes its arrays unchanged through
 * iso_c_binding.  Every call returns 0 on success or a negative CGO_E_*
 * code; cgo_last_error() holds the message.  A handle is not re-entrant:
 * use one per host thread.  Device memory is owned by the handle.
 */
#ifndef CGOCN_H
#define CGOCN_H
#include <stddef.h>
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

#define CGO_OK 0
#define CGO_E_ARG (-1)
#define CGO_E_HIP (-2)
#define CGO_E_STATE (-3)
#define CGO_E_NOMEM (-4)

#define CGO_FP64 0      /* plain FP64 form */
#define CGO_FP32C 1     /* FP32 advection-diffusion, compensated; FP64 co, chemistry, sums */

typedef struct cgo_engine cgo_engine;

/* Grid metrics as the reference computes them in initialise_goldstein
 * (goldstein.f90:906-1075).  Per-j arrays have maxj+2 entries indexed by the
 * Fortran j (0..maxj+1), per-k arrays maxk+2 entries indexed by the Fortran
 * k (0..maxk+1); entries the reference leaves undefined may hold anything.
 * k1 is k1(0:maxi+1, 0:maxj+1) of goldstein_lib.f90:9 (land where k1 > maxk). */
typedef struct cgo_grid_desc {
    int32_t maxi, maxj, maxk;
    const int32_t *k1;
    double dphi, rdphi;
    const double *rc, *rc2, *cv, *cv2, *rds, *dsv, *rdsv, *ds, *sv;
    const double *dt, *rdz, *rdza, *dza, *ssmax, *dz;
    double ec[5];
    double dsc;     /* depth scale in m (goldstein_lib.f90:50), BIOGEM geometry */
} cgo_grid_desc;

typedef struct cgo_member_params {
    double diff[2];   /* non-dimensional diff(1), diff(2) (goldstein.f90:1406-1407) */
    int32_t diso;     /* isoneutral flux on (goldstein-defaults.nml:56) */
    int32_t carry;    /* CGO_FP32C only: carry the rounding residual (int16 plane) */
} cgo_member_params;

/* BIOGEM parameters (biogem-defaults.nml:233-303, gem-defaults.nml:196-199)
 * and the compact tracer slot (0-based, -1 = absent) of each species. */
typedef struct cgo_bgc_params {
    int32_t slot_DIC, slot_ALK, slot_PO4, slot_O2, slot_DOM_C, slot_DOM_P;
    int32_t caco3;
    double dtyr, k0_PO4, c0_PO4, DOMfrac, DOMlifetime;
    double red_POP_POC, red_POP_PO2, red_PON_ALK, red_POP_PON, red_POC_CaCO3, red_POC_CaCO3_pP;
    double POC_frac2, POC_dfrac2, POC_c0frac2, POC_eL1, POC_eL2;
    double CaCO3_frac2, CaCO3_eL1, CaCO3_eL2;
    double sinkrate, remin_k_O2, remin_c0_O2, carb_tol, dsc;
    int32_t carb_maxit;
    /* air-sea exchange of CO2 (needs DIC and ALK) and O2 with a well-mixed
     * ATCHEM atmosphere (biogem_box.f90:27-299, atchem.f90:63-158, 306-320):
     * on/off, par_gastransfer_a (biogem-defaults.nml:221) and the initial
     * partial pressures in atm (atm_init, atchem-defaults.nml:4) */
    int32_t airsea;
    double gastransfer_a, atm_pCO2, atm_pO2;
} cgo_bgc_params;

typedef struct cgo_diag {
    int64_t steps;       /* completed ocean steps */
    int32_t limps;       /* slope-limited points, summed (tstepo_flux 2596-2602) */
    int32_t carb_fail;   /* failed carbonate solves (gem_carbchem.f90:431-478) */
    int64_t carb_iters;  /* summed fixed-point iterations of successful solves */
    double dmax;         /* max isoneutral vertical diffusion number (2605-2608) */
    double cost;         /* summed convection count (co, 2759) */
    double smean;        /* current mean salinity of the BIOGEM state */
} cgo_diag;

/* lifecycle */
/* Build check of the library against this header (no device call):
 * sizes[0..3] = sizeof cgo_grid_desc, cgo_member_params, cgo_bgc_params,
 * cgo_diag as the library was compiled; stamp = hash of this header at build
 * time.  A binding compares both with its own view before any other call. */
int cgo_abi_info(int64_t *sizes, char *stamp, size_t n);
int cgo_create(int32_t device, cgo_engine **out);
int cgo_destroy(cgo_engine *e);
int cgo_last_error(const cgo_engine *e, char *buf, size_t n);
int cgo_sync(cgo_engine *e);

/* grids and members */
int cgo_grid_create(cgo_engine *e, const cgo_grid_desc *g, int32_t *grid_id);   /* maxi <= 143 */
int cgo_member_create(cgo_engine *e, int32_t grid_id, int32_t maxl, int32_t precision,
                      const cgo_member_params *p, int32_t *member_id);
int cgo_member_destroy(cgo_engine *e, int32_t member_id);
int64_t cgo_member_bytes(cgo_engine *e, int32_t grid_id, int32_t maxl, int32_t precision, int32_t bgc);

/* uploads / downloads in reference layouts */
int cgo_set_velocity(cgo_engine *e, int32_t m, const double *u);        /* u(3,0:maxi,0:maxj,maxk) */
int cgo_set_sflux(cgo_engine *e, int32_t m, const double *sflux);       /* ts(:,1:maxi,1:maxj,maxk+1) or NULL */
/* go_ts(maxl,maxi,maxj,maxk); before the member's first step the upload also
 * forms rho from it (initialise_goldstein, goldstein.f90:1450), afterwards
 * rho stays the last tstepo's (step_goldstein's ts = go_ts, 176-177) */
int cgo_upload_state(cgo_engine *e, int32_t m, const double *go_ts);
int cgo_download_state(cgo_engine *e, int32_t m, double *go_ts);
int cgo_download_rho(cgo_engine *e, int32_t m, double *rho);            /* rho(1:maxi,1:maxj,1:maxk) */
int cgo_inventory(cgo_engine *e, int32_t m, double *inv);               /* maxl, diag2 weights */
/* cgo_inventory without any device allocation: no counterpart in the
 * reference; for callers that cannot allocate device memory (tight serving
 * loops, harnesses that forbid hipMalloc).  The sums go through a static
 * device array of 64 doubles per GPU (members with more tracers get
 * CGO_E_ARG); concurrent calls on one GPU, from any engine, share that array
 * and race. */
int cgo_inventory_noalloc(cgo_engine *e, int32_t m, double *inv);
int cgo_get_diag(cgo_engine *e, int32_t m, cgo_diag *d);
/* the convection count cost(maxi,maxj) that co accumulates (goldstein.f90:2759);
 * step_goldstein takes it in from go_cost and hands it back (goldstein.f90:99, 447) */
int cgo_set_cost(cgo_engine *e, int32_t m, const double *cost);
int cgo_download_cost(cgo_engine *e, int32_t m, double *cost);

/* Exact checkpoint of one member (no counterpart in the reference, whose
 * restarts are the netCDF files of goldstein_data.f90:153-354 and
 * biogem.f90:2280-2360; cgenie_amd/restart.py writes and reads those): a
 * header (grid fingerprint, tracer count, precision, carry, parameters,
 * BIOGEM state flags) and every device array of the member (with those of
 * its velocity solve and surface fluxes, where it has them).  cgo_restore
 * into a member of the same grid, tracer count, precision and carry (BIOGEM
 * arrays are added if the checkpoint has them) continues the run bit for
 * bit; anything else is refused with CGO_E_ARG. */
int64_t cgo_checkpoint_bytes(cgo_engine *e, int32_t m);
int cgo_checkpoint(cgo_engine *e, int32_t m, void *buf, size_t n);
int cgo_restore(cgo_engine *e, int32_t m, const void *buf, size_t n);

/* BIOGEM: state from the current go_ts (ocn = ts*S/Smean, T+273.15,
 * S+34.9), sub_init_phys_ocn geometry, [H] start 10**(-7.8).
 * light = solfor/solar_constant (maxi,maxj); mld in m (maxi,maxj). */
void cgo_bgc_default_params(cgo_bgc_params *p, double dtyr);
int cgo_bgc_init(cgo_engine *e, int32_t m, const cgo_bgc_params *p, const double *light, const double *mld);
/* what: 0 = ocn (maxl,maxi,maxj,maxk: T, S and the carried species, other
 * slots zero), 1 = bio_part (5,maxi,maxj,maxk), 2 = carb (10,maxi,maxj),
 * 3 = last anomaly vdocn (maxl,maxi,maxj,maxk), 4 = carbconst (17,maxi,maxj)
 * of the surface cells from the current ocn T and S (sub_calc_carbconst,
 * gem_carbchem.f90:59-268; land columns zero), 5 = bio_settle at k1
 * (5,maxi,maxj), the last step's rain onto the sea floor in mol per step
 * (POC, POP, CaCO3) and the recalcitrant fractions (biogem.f90:1746-1754),
 * 6 = the atmosphere (2,maxi,maxj): pCO2 and pO2 in atm over every column
 * (ATCHEM's atm(ia_pCO2/ia_pO2,:,:), handed to BIOGEM as dum_sfcatm1),
 * 7 = the last step's ocean -> atmosphere flux (2,maxi,maxj) of CO2 and O2
 * in mol yr-1 (locij_focnatm, biogem.f90:1605; land zero) */
int cgo_bgc_download(cgo_engine *e, int32_t m, int32_t what, double *out);
/* wind speed (maxi,maxj) in m/s for the air-sea exchange: the prescribed
 * field par_phys_windspeed that biogem_climate copies to ipoa_wspeed
 * (biogem.f90:2228-2232); zero until set, kept across cgo_bgc_init */
int cgo_set_windspeed(cgo_engine *e, int32_t m, const double *wspeed);

/* GOLDSTEIN's velocity solve (step_goldstein, goldstein.f90:198-233): the
 * frictional-geostrophic barotropic streamfunction with wind forcing, the
 * bottom-pressure torque (JEBAR) and island circulations, then the
 * baroclinic velocities and w from continuity (wind 3685-3714, jbar
 * 3217-3315, ubarsolv 3500-3565, island goldstein_lib.f90:186-241, matmult
 * 3470-3492, velc 3568-3679).  cgo_grid_set_momentum forms, once per grid,
 * what initialise_goldstein does for it (goldstein.f90:1359-1575, 1788-1814:
 * depths, drag, the psi equation's band elimination, the island basis);
 * cgo_set_wind then switches a member from prescribed to solved velocities,
 * and every cgo_step / cgo_tstepo / captured graph solves them before the
 * member's tstepo.  cgo_set_velocity switches a member back to prescribed
 * velocities.  maxi <= 62; the solve's LDS, 8 (maxi (maxj+1) + 96 maxi + 432)
 * bytes, at most 64 KiB (36 x 36: 41 KiB). */
typedef struct cgo_momentum_desc {
    /* per-j metrics the grid descriptor does not carry (maxj+2 entries,
     * Fortran j; goldstein.f90:928-960): s and c, sine and cosine of
     * latitude at T points; rcv = 1/cv; rds2 = 2/(dsv(j)+dsv(j-1)) */
    const double *s, *c, *rcv, *rds2;
    /* the .psiles file (goldstein.f90:1500-1517): the landmass number of
     * every psi point, (maxi, 0:maxj), 0 over the ocean */
    const int32_t *labels;
    int32_t isles;            /* islands: landmasses other than the first (1518) */
    const int32_t *npi;       /* (isles) path lengths, the .paths file (1538-1574) */
    const int32_t *paths;     /* (3, sum npi): lpisl, ipisl, jpisl of each path point, island by island */
    double adrag;             /* inverse minimum drag in days (ini_gold_nml, goin_GOLD.dat: 2.5) */
    double rel;               /* velocity relaxation (ini_gold_nml rel: 0.9) */
} cgo_momentum_desc;
int cgo_grid_set_momentum(cgo_engine *e, int32_t grid_id, const cgo_momentum_desc *d);
/* the wind stress of member m, step_goldstein's stressxu_ocn, stressyu_ocn,
 * stressxv_ocn, stressyv_ocn (maxi, maxj) in N m-2 and scf (ini_gold_nml:
 * 2.0), goldstein.f90:102-107; the first call allocates the member's arrays
 * of the solve and switches it on */
int cgo_set_wind(cgo_engine *e, int32_t m, const double *stressxu, const double *stressyu, const double *stressxv,
                 const double *stressyv, double scf);
/* go_u(3, maxi, maxj, maxk) and go_psi(0:maxi, 0:maxj) of step_goldstein
 * (goldstein.f90:35-37): the velocities of the last step (FP64 members, or
 * any member with a velocity solve) and the barotropic streamfunction */
int cgo_download_velocity(cgo_engine *e, int32_t m, double *go_u);
int cgo_download_psi(cgo_engine *e, int32_t m, double *go_psi);

/* EMBM's surface fluxes (surflux, embm.f90:2548-3738) in the configuration
 * without ENTS, and the surface slot ts(1:2, 1:maxi, 1:maxj, maxk+1) that
 * step_goldstein forms from them (goldstein.f90:149-170).  Per grid, the
 * tables and constants initialise_embm leaves (embm.f90:198-1530);
 * per member, the air temperature and humidity (EMBM's tq), the sea-ice
 * height, cover and surface temperature, and the sea-ice model's heat and
 * freshwater into the ocean.  The ocean side is the member's own top level
 * as its last tstepo left it (tsval_ocn, ssval_ocn, step_goldstein 427-430).
 * Greenhouse gases: compounding at the namelist rates (embm.f90:2916-2926)
 * or, with radfor_atchem, CO2 from the member's ATCHEM atmosphere
 * (2870-2887; CH4 and N2O are not carried and read as their references).
 * Not covered: ENTS, orography, ice sheets, time-series and orbital
 * forcing, freshwater hosing and anomaly files. */
typedef struct cgo_surflux_desc {
    const double *solfor;     /* solfor(maxj, nyear), W m-2 (radfor, embm.f90:2383-2522) */
    int32_t nyear;            /* ocean steps per year: the call count picks solfor(:, mod(istep-1, nyear)+1) */
    const double *albcl;      /* (maxi, maxj) climatological albedo (911-920) */
    const double *ca;         /* (maxi, maxj) shortwave absorbed in the atmosphere (924-929) */
    const double *pmeadj;     /* (maxi, maxj) freshwater adjustment, m s-1 (1292-1345), or NULL for zero */
    const int32_t *iroff, *jroff; /* (maxi, maxj) runoff destination of every cell, 1-based (readroff 3779-3836) */
    int32_t igrid;            /* 0: runoff added as is; else weighted ds(j)/ds(jroff) (3350-3357) */
    int32_t polar_avg;        /* par_wind_polar_avg: 2 leaves the polar rows' wind speed unaveraged */
    int32_t radfor_atchem;    /* atchem_radfor = 'y' */
    double scf, rmax, delf2x; /* embm-defaults.nml: 2.0, 0.85, 5.77 */
    double rdtdim;            /* 1 / (ocean step in s) (770) */
    double rsictscsf;         /* ocean-ice heat transfer (1133) */
    double olr_adj, olr_adj0, t_eqm;      /* 0, 0, 12.371 */
    double albsic_min, albsic_max, sich_max; /* 0.2, 0.7, 9999.9 */
    double rate_co2, rate_ch4, rate_n2o;  /* compound rise per call (761-763) */
    double co2, ch4, n2o;     /* initial mixing ratios (radfor_scl_* x 278e-6, 700e-9, 275e-9) */
} cgo_surflux_desc;
int cgo_grid_set_surflux(cgo_engine *e, int32_t grid_id, const cgo_surflux_desc *d);
/* switches member m's surface fluxes on (atemp non-NULL) with this state,
 * each (maxi, maxj): air temperature (deg C) and specific humidity (kg/kg),
 * sea-ice height (m), cover (0-1) and surface temperature (deg C), the
 * sea-ice model's heat (W m-2) and freshwater (mm s-1) into the ocean; NULL
 * for any after atemp is zero.  From then on every cgo_step and captured
 * graph computes the member's surface slot before its ocean step, and
 * cgo_set_sflux is refused.  atemp NULL switches them off. */
int cgo_set_surface(cgo_engine *e, int32_t m, const double *atemp, const double *ashum, const double *sich,
                    const double *sica, const double *tice, const double *fx0sic, const double *fwsic);
/* one surflux call for the listed members (a host that drives the genie.f90
 * order itself calls it where genie calls surflux_wrapper, genie.f90:271-282,
 * before cgo_tstepo) */
int cgo_surflux(cgo_engine *e, const int32_t *members, int32_t n);
/* the surface slot of member m formed again from its last call's outputs
 * with new sea-ice heat and freshwater fluxes (fx0sic_ocn, fwsic_ocn of
 * step_goldstein, goldstein.f90:149-170; NULL = zero), for a host whose
 * sea-ice model runs between surflux and step_goldstein (genie.f90:271-313) */
int cgo_surface_slot(cgo_engine *e, int32_t m, const double *fx0sic, const double *fwsic);
/* The outputs of member m's last call, CGO_SF_N fields (maxi, maxj) in the
 * order below: surflux's ocean and atmosphere fluxes (W m-2, mm s-1), the
 * sea-ice changes dthsic, dtareasic, the albedos, and EMBM's own fxsw, fx0a,
 * fx0o, fxsen, fxlw, evap, pptn, fx0sic (zero without ice), fx0neto and
 * tsfreez; ashum and tice (maxi, maxj) as the call left them; gases the
 * CO2, CH4, N2O mixing ratios and the call count.  Any pointer may be NULL. */
#define CGO_SF_N 28
enum { CGO_SF_FXLHO = 0, CGO_SF_FXSHO, CGO_SF_FXSWO, CGO_SF_FXLWO, CGO_SF_EVAP_OCN, CGO_SF_PPTN_OCN,
       CGO_SF_RUNOFF_OCN, CGO_SF_RUNOFF_LAND, CGO_SF_FXLHA, CGO_SF_FXSHA, CGO_SF_FXSWA, CGO_SF_FXLWA,
       CGO_SF_EVAP_ATM, CGO_SF_PPTN_ATM, CGO_SF_DTHSIC, CGO_SF_DTAREASIC, CGO_SF_ALBEDO, CGO_SF_ALBICE,
       CGO_SF_FXSW, CGO_SF_FX0A, CGO_SF_FX0O, CGO_SF_FXSEN, CGO_SF_FXLW, CGO_SF_EVAP, CGO_SF_PPTN,
       CGO_SF_FX0SIC, CGO_SF_FX0NETO, CGO_SF_TSFREEZ };
int cgo_download_surflux(cgo_engine *e, int32_t m, double *out, double *ashum, double *tice, double *gases);

/* stepping: nsteps ocean steps of every listed member, batched into one
 * launch per kernel (members with BIOGEM run the full coupling order:
 * cgo_tstepo, cgo_bio_step, cgo_tracercoupling) */
int cgo_step(cgo_engine *e, const int32_t *members, int32_t n, int32_t nsteps);

/* The three calls of one ocean step for a host that drives the genie.f90
 * order itself (genie.f90:306-313, 360-381).
 *
 * cgo_tstepo replaces tstepo (goldstein.f90:2280-2432) as called from
 * step_goldstein (goldstein.f90:17-38, 235): tstepo_flux + eos, co, and the
 * ts1 = ts hand-over, nsteps times, each preceded by the velocity solve for
 * the members that have one (198-233).
 * cgo_bio_step replaces step_biogem (biogem.f90:528-547) for the biology,
 * remineralisation and surface chemistry of the carried species; dtyr is the
 * one given to cgo_bgc_init.
 * cgo_tracercoupling replaces biogem_tracercoupling(dum_ts, dum_ts1)
 * (biogem.f90:1885-1890); go_ts and go_ts1 are the one device state.  As in
 * the reference it must follow cgo_bio_step (it applies the anomalies that
 * call leaves), and nothing else may step those members in between
 * (CGO_E_STATE otherwise). */
int cgo_tstepo(cgo_engine *e, const int32_t *members, int32_t n, int32_t nsteps);
int cgo_bio_step(cgo_engine *e, const int32_t *members, int32_t n);
int cgo_tracercoupling(cgo_engine *e, const int32_t *members, int32_t n);
/* nsteps of cgo_step captured once as a HIP graph and replayed */
int cgo_graph_capture(cgo_engine *e, const int32_t *members, int32_t n, int32_t nsteps, int32_t *graph_id);
int cgo_graph_launch(cgo_engine *e, int32_t graph_id, int32_t nreps);
int cgo_graph_destroy(cgo_engine *e, int32_t graph_id);

/* tuning and measurement */
int cgo_set_tile(cgo_engine *e, int32_t precision, int32_t tracer_block);   /* 2, 4, 6 or 8 */
int cgo_set_flux_cols(cgo_engine *e, int32_t precision, int32_t cols);      /* 128 or 256 wet columns per block */
int cgo_set_flux_ring(cgo_engine *e, int32_t precision, int32_t planes);    /* 2 or 3 level planes in LDS */
int cgo_set_flux_iso(cgo_engine *e, int32_t precision, int32_t prepass);    /* 1: isoneutral coefficients from a pre-pass */
/* average device time (ms) of one launch of kernel `which` (0 = flux,
 * 1 = co, 2 = bgc column + coupling, 3 = flux with the BIOGEM column step
 * running beside it on the second stream, as inside a whole step, 4 = the
 * BIOGEM column step alone, 5 = the velocity solve, 6 = the surface fluxes) over the listed members, from HIP events on the
 * engine stream; the state is advanced by the timed launches (mode 4
 * rewrites the BIOGEM arrays at every launch with no coupling between them:
 * a member timed with it is for timing only) */
int cgo_time_kernel(cgo_engine *e, int32_t which, const int32_t *members, int32_t n, int32_t reps, double *avg_ms);

#ifdef __cplusplus
}
#endif
#endif
