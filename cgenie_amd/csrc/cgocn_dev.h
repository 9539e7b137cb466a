// cgocn_dev.h -- device-side data model of the MI355X ocean tracer engine.
//
// Layout in HBM (per member, i fastest, no halos):
//   state  [L][maxk][maxj][maxi]   (double, or float + int16 carry plane)
//   u      [3][maxk][maxj][maxi]   u(n,i,j,k) of goldstein.f90:684 (interior)
//   sflux  [L][maxj][maxi]         ts(l,i,j,maxk+1), goldstein.f90:167-170
//   rho    [maxk][maxj][maxi]
// Per grid: k1 with its reference halo [(maxj+2)][(maxi+2)] (land rows at
// j = 0 and maxj+1, periodic copies at i = 0 and maxi+1, goldstein.f90:1114-1123)
// and per-j / per-k metric tables.  Fortran indices are used in the kernels
// where they index the metric tables (I = i0 + 1, J = j0 + 1, k = kz + 1).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace cgo {

// per-j metric table rows, each of length maxj+2 (index = Fortran j)
enum { MJ_RC = 0, MJ_RC2, MJ_CV, MJ_CV2, MJ_RDS, MJ_DSVM, MJ_RDSV, MJ_DS, MJ_SV, NMJ };
// per-k metric table rows, each of length maxk+2 (index = Fortran k)
enum { MK_DT = 0, MK_RDZ, MK_RDZA, MK_DZA, MK_SSMAX, MK_DZ, MK_DBOT, MK_DMID, MK_DD, NMK };

// BIOGEM sediment (particulate) slots in the reference's is order
enum { S_POC = 0, S_POP, S_CACO3, S_POCFR2, S_CACO3FR2, NSED };
// carb descriptors (gem_cmn.f90:268-277, 0-based)
enum { IC_H = 0, IC_FUG_CO2, IC_CONC_CO2, IC_CONC_CO3, IC_CONC_HCO3, IC_OHM_CAL, IC_OHM_ARG,
       IC_DCO3_CAL, IC_DCO3_ARG, IC_RF0, NCARB };
// the BGC species the engine carries, in the reference's io order
// (io_DIC=3, io_PO4=8, io_O2=10, io_ALK=12, io_DOM_C=15, io_DOM_P=20)
enum { B_DIC = 0, B_PO4, B_O2, B_ALK, B_DOMC, B_DOMP, NBGC };

constexpr int MAXK_DEV = 48;      // deepest grid the column kernels accept
constexpr double NULLSMALL = +0.999999E-19;  // gem_cmn.f90:719
constexpr double ZEROC = 273.15;
constexpr double SALN0 = 34.9;
// unit conversions and constants of gem_cmn.f90:510-545, 690-698
constexpr double CONV_YR_HR = 24.0 * 365.25;
constexpr double CONV_YR_S = 3600.0 * (24.0 * 365.25);
constexpr double CONV_S_YR = 1.0 / (3600.0 * (24.0 * 365.25));
constexpr double CONV_CM_M = 1.0 / 100.0;
constexpr double CONV_PA_ATM = 1.0 / 1.01325E+05;
constexpr double CONST_R_SI = 8.3145;
constexpr double CONST_V = 0.022414;
constexpr double CONST_PI = 3.141592653589793;
constexpr double CONST_REARTH = 6.37E+06;
constexpr double ATM_TH = 7777.0;  // par_atm_th, atchem_lib.f90:86

// One block of the flux kernel: wet columns [w0, w1) of the ascending order,
// staged (with every neighbour they read) over flattened columns [t0, t1).
struct FluxBlk {
    int32_t w0, w1, t0, t1;
};
// tile positions staged per thread of a flux block of nt threads
__host__ __device__ constexpr int flux_pf(int nt) { return nt == 128 ? 3 : 2; }
// positions of one LDS slot of a flux block (<= flux_pf(nt) * nt): the host
// cuts blocks so that every staged range fits, and the fixed stride makes
// each slot's offset an immediate of the LDS instruction.  Close to the
// widest range of a 36x36 mask (266 and 439 positions) so the slots cost no
// resident blocks; a single column spans 2 maxi + 1 positions.
__host__ __device__ constexpr int flux_tw(int nt) { return nt == 128 ? 288 : 448; }
// the flux grid of a batch: blocks and widest tile, per block width (128, 256)
struct FluxShape {
    int nblk[2], tw[2];
};

struct GridDev {
    int maxi, maxj, maxk, ncol, ncell, nwet;
    const FluxBlk* fblk[2];  // flux blocks for 128- and 256-thread workgroups
    int nfblk[2];
    const int32_t* k1;      // [(maxj+2)*(maxi+2)]
    const double* mj;       // [NMJ*(maxj+2)]
    const double* mk;       // [NMK*(maxk+2)]
    const float* mjf;       // float copies for the FP32 form
    const float* mkf;
    const int32_t* wetcols; // [nwet] column ids i0 + maxi*j0, reference vector order (i outer, j inner)
    const int32_t* wetsorted; // [nwet] the same ids ascending (i fastest): the thread order of column kernels
    const int32_t* vpos;    // [nwet] reference vector position of wetsorted[w] (partial-sum slot)
    const int32_t* cslot;   // [ncol] partial-sum slot of each column, -1 for land
    const double* V;        // [ncell] cell volumes (biogem_data.f90:1128)
    double dphi, rdphi, rVtot;
    double ec[5];
};

struct BgcParams {
    int32_t slot[NBGC];     // compact tracer slot of each species, -1 if absent
    int32_t ocn[NBGC];      // plane of each species in the ocn block, -1 if absent
    int32_t caco3;
    double dtyr, k0_PO4, c0_PO4, DOMfrac, DOMlifetime;
    double red_POP_POC, red_POP_PO2, red_PON_ALK, red_POP_PON, red_POC_CaCO3, red_POC_CaCO3_pP;
    double POC_frac2, POC_dfrac2, POC_c0frac2, POC_eL1, POC_eL2;
    double CaCO3_frac2, CaCO3_eL1, CaCO3_eL2;
    double sinkrate, remin_k_O2, remin_c0_O2, carb_tol, dsc;
    int32_t carb_maxit;
    int32_t airsea;         // CO2 and O2 exchange with the well-mixed atmosphere
    double gas_a;           // par_gastransfer_a
};

// GOLDSTEIN's velocity solve of one grid (goldstein.f90:198-233): the
// factors formed once on the host (the psi equation's band elimination, the
// island basis, depths, drag) and the metric tables the solve reads.
// Fortran indices throughout; layouts (l fastest where a leading l is given):
//   rh    (3, 0:maxi+1, 0:maxj+1)     drag (2, 0:maxi+1, 0:maxj+1)
//   rtv, rtv3 (maxi, maxj)            ku   (2, maxi, maxj)
//   mk    (0:maxi+1, 0:maxj+1)        getj (maxi, maxj) bytes
//   fwd   [N-1][n+1]: row i, entry q-1 = ratm(i+q, q), the multiplier that
//         elimination step i applies to row i+q (0 past row N)
//   bwd   [N][n+2]:   row i, entry k = gap(i, n+2+k) of the eliminated band
//   bmask [N]: row i, bit k-1 set where gap(i, n+2+k), k >= 1, is not zero
//   ubisl isles x (2, 0:maxi+1, 0:maxj), psisl isles x (0:maxi, 0:maxj)
//   erisl (isles, isles+1) column-major, inverted in place (matinv_gold)
//   paths (mpi, isles): lpisl, ipisl, jpisl; npi (isles)
//   ainv  [N][N] row-major: the psi equation's inverse (the FP32 form's
//         dense solve), or null
enum { JT_S = 0, JT_SV, JT_C, JT_CV, JT_RC, JT_RCV, JT_RDS, JT_RDSV, JT_RDS2, JT_DS, JT_DSV, JT_N };
struct MomGridDev {
    int maxi, maxj, maxk, n, N, isles, mpi;
    double dphi, rdphi, rel;
    const double* jt;       // [JT_N][maxj+2]
    const double* dz;       // [maxk+2]
    const double* dza;      // [maxk+2]
    const double* rh;
    const double* drag;
    const double* rtv;
    const double* rtv3;
    const int32_t* ku;
    const int32_t* mk;
    const uint8_t* getj;
    const double* fwd;
    const double* bwd;
    const uint64_t* bmask;
    const double* ubisl;
    const double* psisl;
    const double* erisl;
    const int32_t* npi;
    const int32_t* lpisl;
    const int32_t* ipisl;
    const int32_t* jpisl;
    const double* ainv;
};
constexpr int MOM_MAXISL = 16;   // islands
constexpr int MOM_DENSE_MAXN = 4096;  // psi points up to which a grid keeps the dense inverse
constexpr int MOM_SPLITK = 8;         // slices of the dense solve's sums

// one member's arrays of the velocity solve, offsets in doubles from its base
struct MomMemberLayout {
    size_t gbw, tau, dztau, dztav, u1, ub, psi, bp, sbp, csum, psibc, u64, gst, total;
};
__host__ __device__ inline MomMemberLayout mom_member_layout(int mi, int mj, int mk, bool own_u64) {
    MomMemberLayout L{};
    const size_t ncol = (size_t)mi * mj, ncell = ncol * mk, N = (size_t)mi * (mj + 1);
    size_t o = 0;
    auto take = [&](size_t n) {
        size_t r = o;
        o += (n + 31) & ~(size_t)31;
        return r;
    };
    L.gbw = take(N);                                       // the wind part of the psi forcing (wind, 3685-3714)
    L.tau = take(2 * ncol);                                // tau (2, maxi, maxj)
    L.dztau = take(2 * ncol);                              // dztau (2, maxi, maxj)
    L.dztav = take(2 * ncol);                              // dztav (2, maxi, maxj)
    L.u1 = take(2 * ncell);                                // u1 of u and v, [2][maxk][maxj][maxi]
    L.ub = take((size_t)2 * (mi + 2) * (mj + 1));          // ub (2, 0:maxi+1, 0:maxj)
    L.psi = take((size_t)(mi + 1) * (mj + 1));             // psi (0:maxi, 0:maxj)
    L.bp = take((size_t)(mi + 2) * (mj + 2) * (mk + 1));   // bp (0:maxi+1, 0:maxj+1, 0:maxk)
    L.sbp = take((size_t)(mi + 2) * (mj + 2));             // sbp (0:maxi+1, 0:maxj+1)
    L.csum = take(2 * ncol);                               // velc's depth sums of the shear, [2][maxj][maxi]
    L.psibc = take(MOM_MAXISL);                            // the islands' streamfunction values of the step
    L.u64 = own_u64 ? take(3 * ncell) : 0;                 // FP64 u, v, w of an FP32 member
    L.gst = own_u64 ? take((1 + MOM_SPLITK) * N) : 0;      // an FP32 member's psi forcing, then psi's partial sums (the dense solve)
    L.total = o;
    return L;
}

// EMBM's surface fluxes of one grid (surflux, embm.f90:2548-3738, without
// ENTS): the set-up tables and the namelist constants, and the runoff
// routing turned into gather lists, each wet column's land sources in the
// order the reference's main loop (i outer, j inner) adds them.
struct SfxGridDev {
    int32_t nyear, igrid, radfor_atchem, nsrc;
    double rmax, delf2x, rdtdim, rsictscsf, olr_adj, olr_adj0, t_eqm;
    double albsic_min, albsic_max, sich_max, rate_co2, rate_ch4, rate_n2o;
    const double* solfor;   // [nyear][maxj]
    const double* albcl;    // [ncol]
    const double* ca;       // [ncol]
    const double* pmeadj;   // [ncol]
    const int32_t* rstart;  // [ncol + 1] range of each column's sources in rsrc
    const int32_t* rsrc;    // [nsrc] source columns
};
// per-member fields of the surface fluxes, each [ncol]: the air and sea-ice
// state (ashum and tice in/out), the wind speed, the sea-ice model's heat
// and freshwater into the ocean, and the top level's T and S as the last
// tstepo left them (tsval_ocn, ssval_ocn of step_goldstein 427-430)
enum { SW_ATEMP = 0, SW_ASHUM, SW_SICH, SW_SICA, SW_TICE, SW_USURF, SW_FX0SIC, SW_FWSIC, SW_TSTAR, SW_SSTAR, SW_NF };
// outputs of one call, each [ncol], in the order of include/cgocn.h's CGO_SF_*
enum {
    SF_FXLHO = 0, SF_FXSHO, SF_FXSWO, SF_FXLWO, SF_EVAP_OCN, SF_PPTN_OCN, SF_RUNOFF_OCN, SF_RUNOFF_LAND, SF_FXLHA,
    SF_FXSHA, SF_FXSWA, SF_FXLWA, SF_EVAP_ATM, SF_PPTN_ATM, SF_DTHSIC, SF_DTAREASIC, SF_ALBEDO, SF_ALBICE, SF_FXSW,
    SF_FX0A, SF_FX0O, SF_FXSEN, SF_FXLW, SF_EVAP, SF_PPTN, SF_FX0SIC, SF_FX0NETO, SF_TSFREEZ, SF_N
};
// member scalars after the fields and outputs: the greenhouse gases and the call count
enum { SS_CO2 = 0, SS_CH4, SS_N2O, SS_ISTEP, SS_N };
__host__ __device__ inline size_t sfx_member_doubles(int ncol) { return (size_t)(SW_NF + SF_N) * ncol + SS_N; }

struct MemberDev {
    GridDev g;
    int32_t L, prec, diso, bgc;
    double diff1, diff2;
    float diff1f, diff2f;
    void* st[2];            // ping-pong state buffers
    int16_t* lo[2];         // FP32 form: carry planes (fraction of an ulp / 2^16), or null
    const void* u;          // velocity (double, or float in the FP32 form)
    const double* sflux;    // [L][ncol] or null
    double* rho;            // [ncell]
    double* isoc;           // isoneutral coefficients of the flux pre-pass: [10][ncell] doubles
                            // (dzrho or +1, dxr[4], dyr[4], slim) or [9][ncell] floats (ax[4], ay[4], qq)
    double* cost;           // [ncol]
    int32_t* parity;        // device scalar: which st[] is current
    int64_t* steps;         // device scalar: completed steps
    // BIOGEM
    BgcParams bp;
    double* ocn;            // [2 + nbgc_present][ncell]: T(K), S, BGC species (absolute)
    double* part;           // [NSED][ncell]
    double* settle;         // [NSED][ncol] bio_settle at k1
    double* carb;           // [NCARB][ncol]
    double* M;              // [ncell]
    double* rM;             // [ncell]
    double* vd;             // [NBGC][ncell] anomalies of the present species
    double* light;          // [ncol]
    double* mld;            // [ncol]
    double* wspeed;         // [ncol] wind speed (m/s) of the air-sea exchange
    double* atm;            // [2][ncol] atmospheric pCO2, pO2 (atm) over every column
    double* fatm;           // [2][ncol] the last BIOGEM step's ocean -> atmosphere flux (mol yr-1)
    double* colsum;         // [(L+1)][nwet] column partial sums (NEW_l, S_new)
    double* colold;         // [L][nwet] column partial sums of OLD_l for the next coupling
    double* scal;           // member scalars, see SC_*
    const double* rtab;     // [RT_N][maxk+2] remineralisation factors per level (host-computed)
    const int32_t* mink;    // [maxk+2] deepest-reach level of particles leaving level k (0 = bottom)
    int32_t* diag;          // [4] limps, carb_fail, carb_iters, co guard hits
    int32_t* sfresh;        // device flag: colsum holds the sums of the current state
    double* dmax;           // device scalar
    // velocity solve: null when the velocities are prescribed
    const MomGridDev* mg;
    double* mw;             // the member's arrays of the solve (mom_member_layout)
    double* u64;            // FP64 u, v, w [3][maxk][maxj][maxi] (the u planes of an FP64 member)
    // surface fluxes: null when the surface slot is prescribed
    const SfxGridDev* sg;
    double* sw;             // [SW_NF][ncol] fields, [SF_N][ncol] outputs, [SS_N] scalars
    double* sfw;            // the member's surface slot, sflux planes 0 and 1, which the fluxes write
};

// remineralisation table rows: layer-thickness ratio dD(k+1)/dD(k) and the
// e-folding fractions 1 - exp(-dD/eL) (biogem_box.f90:2555-2561, 2583-2584, 2664-2689)
enum { RT_LAYER = 0, RT_POC1, RT_POC2, RT_CAL1, RT_CAL2, RT_N };

// member scalar slots (scal[])
// SC_SMEAN is the mean salinity the next coupling starts from; the coupling's
// reduction moves it to SC_SOLD and stores the new mean in both SC_SNEW and SC_SMEAN.
// SC_PSCALE: the coupling's salinity ratio for the particulates, applied when
// they are next read (the BIOGEM column step rewrites them all before the
// following coupling), 1 after initialisation.
enum { SC_SMEAN = 0, SC_SNEW, SC_SRATIO, SC_RSRATIO, SC_SOLD, SC_PSCALE, SC_FACT0 };  // SC_FACT0 + l : OLD_l * rNEW_l

// ---------------------------------------------------------------- helpers
// Pointers read from the member table are generic (flat) to the compiler;
// casting them to the global address space turns flat loads/stores into
// global_load/global_store, which do not tie up the LDS counter.
#define AS_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ AS_GLOBAL T* gp(T* p) { return (AS_GLOBAL T*)p; }
template <typename T>
__device__ __forceinline__ const AS_GLOBAL T* gp(const T* p) { return (const AS_GLOBAL T*)p; }

// Element i of a per-member array, addressed as the array's base plus a
// 32-bit byte offset: with a base uniform over the wave the load or store is
// issued as scalar base + vector offset, with no 64-bit address arithmetic per
// lane.  Every per-member array is below 2 GiB (checked at member creation).
template <typename T>
__device__ __forceinline__ const AS_GLOBAL T& at(const AS_GLOBAL T* b, uint32_t i) {
    return *(const AS_GLOBAL T*)((const AS_GLOBAL char*)b + i * (uint32_t)sizeof(T));
}
template <typename T>
__device__ __forceinline__ AS_GLOBAL T& at(AS_GLOBAL T* b, uint32_t i) {
    return *(AS_GLOBAL T*)((AS_GLOBAL char*)b + i * (uint32_t)sizeof(T));
}

// Workgroups of a launch are dealt to the 8 XCDs (each with its own L2)
// round-robin in dispatch order.  xcd_block renumbers the dispatch index h of
// an n-block launch so that every XCD runs one contiguous range of the
// renumbered blocks: blocks that share data (the tiles and tracer blocks of
// one member) then share an L2.  A bijection of [0, n) for any n.
constexpr int NXCD = 8;
__device__ __forceinline__ int xcd_block(int h, int n) {
    const int q = n / NXCD, rem = n % NXCD, x = h % NXCD, r = h / NXCD;
    return x * q + (x < rem ? x : rem) + r;
}
// the dispatch index of this block and the launch's block count
__device__ __forceinline__ int dispatch_index() {
    return (int)(blockIdx.x + gridDim.x * (blockIdx.y + gridDim.y * blockIdx.z));
}
__device__ __forceinline__ int dispatch_count() { return (int)(gridDim.x * gridDim.y * gridDim.z); }
// this block's (x, y, z) after the renumbering (x fastest, as dispatched)
struct BlockXYZ {
    int x, y, z;
};
__device__ __forceinline__ BlockXYZ xcd_blk() {
    const int v = xcd_block(dispatch_index(), dispatch_count());
    const int gx = (int)gridDim.x, gy = (int)gridDim.y;
    return BlockXYZ{v % gx, (v / gx) % gy, v / (gx * gy)};
}

// unbiased binary exponent of a float, zero and subnormals taken as -126
__device__ __forceinline__ int carry_exp(float x) {
    const int eb = (int)((__float_as_uint(x) >> 23) & 0xffu);
    return (eb > 1 ? eb : 1) - 127;
}

__device__ __forceinline__ int k1at(const GridDev& g, int I, int J) { return ((const AS_GLOBAL int32_t*)g.k1)[J * (g.maxi + 2) + I]; }

// eos, goldstein.f90:3048-3061 (ieos = 0)
template <typename R>
__device__ __forceinline__ R eos(R e0, R e1, R e2, R e3, R t, R s) {
    return e0 * t + e1 * s + e2 * (t * t) + e3 * (t * t * t);
}

}  // namespace cgo
