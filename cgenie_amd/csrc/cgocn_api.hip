// cgocn_api.hip -- host side of libcgocn.so: the C ABI declared in
// include/cgocn.h.  Handles own a HIP stream, the per-grid constant tables and
// one device allocation per member; kernels are launched batched over member
// lists (one launch per kernel per step covers every listed member).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/cgocn.h"
#include "cgocn_dev.h"
#include "mom_host.h"

namespace cgo {
void launch_flux64(const MemberDev*, const int32_t*, int nlist, const FluxShape&, int maxL, int lb, int nt, int nbuf,
                   int iso_split, int part, size_t lds_cap, hipStream_t);
void launch_flux32(const MemberDev*, const int32_t*, int nlist, const FluxShape&, int maxL, int lb, int nt, int nbuf,
                   int iso_split, int part, size_t lds_cap, hipStream_t);
void launch_co(const MemberDev*, const int32_t*, int nlist, int nwetmax, int maxkmax, int prec, hipStream_t);
void launch_flip(const MemberDev*, const int32_t*, int nlist, hipStream_t);
void launch_bio(const MemberDev*, const int32_t*, int nlist, int nwetmax, int W, int prec, hipStream_t);
void launch_coupling(const MemberDev*, const int32_t*, int nlist, int nwetmax, int maxkmax, int maxL, int prec,
                     size_t lds_cap, int phase, hipStream_t);
void launch_inventory(const MemberDev*, int32_t member, int L, double* out, const double* w, int prec, hipStream_t);
void launch_carbconst_surface(const MemberDev*, int32_t member, int nwet, double* out, hipStream_t);
void launch_atchem(const MemberDev*, const int32_t*, int nlist, int ncolmax, hipStream_t);
void launch_mom(const MemberDev*, const int32_t*, int nlist, const int32_t* band, int nband, const int32_t* dense,
                const int* groups, int ngroups, int maxN, int maxcol, int maxn, int part, hipStream_t);
size_t mom_solve_lds(int N, int n);
void launch_sfx(const MemberDev*, const int32_t*, int nlist, int maxcol, int slot_only, hipStream_t);
size_t sfx_lds(int ncol);
size_t mom_ring_terms(int n);
}  // namespace cgo

using namespace cgo;

namespace {

constexpr int TABLE_CAP = 4096;
constexpr int NCC_H = 17;  // carbonate constants per cell (n_carbconst, gem_cmn.f90:66)
constexpr double CONV_M3_KG = 1027.649;         // gem_cmn.f90:510

struct GridH {
    cgo_grid_desc d{};
    GridDev dev{};
    std::vector<int32_t> k1;      // halo'd
    std::vector<double> mj, mk;   // tables
    std::vector<int32_t> wet;     // wet column ids, vocn order
    std::vector<int32_t> wets, vpos; // ascending ids and their vocn positions
    std::vector<double> V;        // [ncell]
    std::vector<double> invw;     // [ncell] inventory weights dphi*ds(j)*dz(k)
    int ftw[2] = {0, 0};          // widest flux tile for 128- and 256-thread blocks
    double Vtot = 0;
    uint64_t fingerprint = 0;     // of the dimensions, mask and metric tables (checkpoints)
    void* dbuf = nullptr;
    double* d_invw = nullptr;
    // velocity solve (cgo_grid_set_momentum): the host set-up, its device
    // copy (one allocation holding the MomGridDev struct and its arrays) and
    // a fingerprint of its inputs (checkpoints)
    MomSetup* mom = nullptr;
    void* mbuf = nullptr;
    const MomGridDev* d_mg = nullptr;
    uint64_t mom_fp = 0;
    bool ainv_on = false;         // the dense inverse is kept (FP32 members solve by it)
    // surface fluxes (cgo_grid_set_surflux): the device tables, the settings
    // the host side needs (wind speed, initial gases) and a fingerprint
    void* sbuf = nullptr;
    const SfxGridDev* d_sg = nullptr;
    double sfx_scf = 2.0, sfx_gas[3] = {0, 0, 0};
    int sfx_polar = 0, sfx_atchem = 0;
    uint64_t sfx_fp = 0;
};

struct MemberH {
    bool alive = false;
    int grid = -1, L = 0, prec = 0, bgc = 0;
    int wstream = 4;  // particle-stream window of the BIOGEM march (4, 5, 6, 8 or 16)
    bool coupling_due = false;  // cgo_bio_step ran; cgo_tracercoupling must follow
    // a step, graph or kernel of this member may still be running on the
    // engine stream (set when one is enqueued, cleared when the stream drains)
    bool queued = false;
    // the member has been stepped: rho is tstepo's from now on
    bool stepped = false;
    MemberDev dev{};
    cgo_member_params p{};
    cgo_bgc_params bp{};
    void* dbuf = nullptr;
    size_t bytes = 0;
    // velocity solve on (cgo_set_wind): the member's arrays of the solve
    void* mbuf = nullptr;
    size_t mbytes = 0;
    // the last wind stress (cgo_set_wind), [4][ncol]: sxu, syu, sxv, syv
    std::vector<double> stress;
    // surface fluxes on (cgo_set_surface): the member's fields, outputs and scalars
    void* sbuf = nullptr;
    size_t sbytes = 0;
};

struct Graph {
    hipGraphExec_t exec = nullptr;
    hipGraph_t graph = nullptr;
    int32_t* d_lists = nullptr;
    // the captured members and the generation of their slots at capture:
    // launch extents, precision lists and stream windows are baked into the
    // graph, so a replay is refused once any of them was destroyed,
    // re-created or re-laid-out
    std::vector<int32_t> ids;
    std::vector<uint64_t> gens;
};

// host copy of the device carry weight: 2^-16 ulp(x)
double carry_unit_h(float x) {
    int e = (x == 0.0f) ? -126 : ilogbf(x);
    if (e < -126) e = -126;
    return ldexp(1.0, e - 23 - 16);
}

}  // namespace

struct cgo_engine {
    int device = 0;
    hipStream_t s = nullptr;
    hipStream_t s2 = nullptr;              // BIOGEM columns, concurrent with transport
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::string err;
    std::vector<GridH*> grids;
    std::vector<MemberH*> members;
    // device buffers given up while other members may be stepping: hipFree
    // waits for the whole device, so they are freed when the stream drains
    std::vector<void*> graveyard;
    MemberDev* d_table = nullptr;
    bool table_dirty = false;
    // per member slot: bumped whenever the slot's member is created,
    // destroyed or re-laid-out (BIOGEM state added), so captured graphs can
    // tell that a slot no longer holds what they captured
    std::vector<uint64_t> gen;
    // flux tile per precision (FP64, FP32C): the shapes the autotuner picks
    // for a large ensemble on MI355X; cgo_set_tile and its companions change them
    int lb[2] = {8, 6};       // tracers per block
    int nt[2] = {256, 256};   // wet columns per block
    int nbuf[2] = {2, 2};     // level planes resident in LDS
    int isosplit[2] = {1, 1}; // isoneutral coefficients from a pre-pass
    size_t lds_cap = 65536;   // LDS one workgroup may take
    std::vector<Graph> graphs;
    int32_t* d_list = nullptr;  // scratch lists for eager steps
    size_t d_list_cap = 0;
    std::vector<int32_t> last_lists;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
};

// Device memory of cgo_inventory_noalloc: the per-tracer sums of one call,
// in place of the buffer cgo_inventory allocates.  Every GPU the library
// runs on holds its own copy until the process ends.  Host code never uses
// the array's name as an address: the device address of each GPU's copy is
// taken once per device, with that device current, by
// hipGetSymbolAddress (noalloc_setup, called from cgo_create), and only
// that address reaches kernels and copies.
constexpr int NOALLOC_MAXL = 64;  // tracers one allocation-free inventory takes
static __device__ __align__(256) double scratch[NOALLOC_MAXL];
constexpr int NOALLOC_MAXDEV = 64;
static double* noalloc_ptr[NOALLOC_MAXDEV] = {};
static std::mutex noalloc_mu;

static int fail(cgo_engine* e, int code, const char* fmt, ...) {
    char buf[512];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    if (e) e->err = buf;
    return code;
}

#define HIPCHK(e, x)                                                                     \
    do {                                                                                 \
        hipError_t _r = (x);                                                             \
        if (_r != hipSuccess) return fail(e, CGO_E_HIP, "%s: %s", #x, hipGetErrorString(_r)); \
    } while (0)

static MemberH* member(cgo_engine* e, int32_t m) {
    if (!e || m < 0 || m >= (int32_t)e->members.size() || !e->members[m] || !e->members[m]->alive) return nullptr;
    return e->members[m];
}

static void bump_gen(cgo_engine* e, int32_t slot) {
    if ((int32_t)e->gen.size() <= slot) e->gen.resize(slot + 1, 0);
    e->gen[slot]++;
}

static int sync_table(cgo_engine* e) {
    if (!e->table_dirty) return CGO_OK;
    // work already enqueued (an eager step or a graph replay) reads the
    // table: let it finish before the entries are rewritten
    HIPCHK(e, hipStreamSynchronize(e->s));
    std::vector<MemberDev> t(e->members.size());
    for (size_t i = 0; i < e->members.size(); i++)
        if (e->members[i] && e->members[i]->alive) t[i] = e->members[i]->dev;
    if (!t.empty()) HIPCHK(e, hipMemcpy(e->d_table, t.data(), t.size() * sizeof(MemberDev), hipMemcpyHostToDevice));
    e->table_dirty = false;
    return CGO_OK;
}

#ifndef CGO_ABI_STAMP
#error "CGO_ABI_STAMP (hash of include/cgocn.h) is set by cgenie_amd/build.py"
#endif

extern "C" {

int cgo_abi_info(int64_t* sizes, char* stamp, size_t n) {
    if (!sizes || !stamp || n == 0) return CGO_E_ARG;
    sizes[0] = (int64_t)sizeof(cgo_grid_desc);
    sizes[1] = (int64_t)sizeof(cgo_member_params);
    sizes[2] = (int64_t)sizeof(cgo_bgc_params);
    sizes[3] = (int64_t)sizeof(cgo_diag);
    snprintf(stamp, n, "%s", CGO_ABI_STAMP);
    return CGO_OK;
}

// the device address of this device's scratch array, once per device
// (the device is current); false if the runtime cannot give it
static bool noalloc_setup(int device) {
    if (device >= NOALLOC_MAXDEV) return true;  // such devices just have no allocation-free path
    std::lock_guard<std::mutex> lk(noalloc_mu);
    if (noalloc_ptr[device]) return true;
    double* ptr = nullptr;
    if (hipGetSymbolAddress((void**)&ptr, HIP_SYMBOL(scratch)) != hipSuccess || !ptr) return false;
    noalloc_ptr[device] = ptr;
    return true;
}

int cgo_create(int32_t device, cgo_engine** out) {
    if (!out) return CGO_E_ARG;
    *out = nullptr;
    int n = 0;
    if (hipGetDeviceCount(&n) != hipSuccess || n <= device || device < 0) return CGO_E_HIP;
    cgo_engine* e = new cgo_engine();
    e->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&e->s, hipStreamNonBlocking) != hipSuccess ||
        hipStreamCreateWithFlags(&e->s2, hipStreamNonBlocking) != hipSuccess ||
        hipMalloc(&e->d_table, TABLE_CAP * sizeof(MemberDev)) != hipSuccess ||
        hipEventCreate(&e->ev0) != hipSuccess || hipEventCreate(&e->ev1) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&e->ev_join, hipEventDisableTiming) != hipSuccess || !noalloc_setup(device)) {
        delete e;
        return CGO_E_HIP;
    }
    int lds = 0;
    if (hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, device) == hipSuccess && lds > 0)
        e->lds_cap = (size_t)lds;
    *out = e;
    return CGO_OK;
}

int cgo_destroy(cgo_engine* e) {
    if (!e) return CGO_E_ARG;
    hipSetDevice(e->device);
    hipStreamSynchronize(e->s);
    for (void* p : e->graveyard) hipFree(p);
    for (auto& g : e->graphs) {
        if (g.exec) hipGraphExecDestroy(g.exec);
        if (g.graph) hipGraphDestroy(g.graph);
        if (g.d_lists) hipFree(g.d_lists);
    }
    for (auto* m : e->members)
        if (m) {
            if (m->dbuf) hipFree(m->dbuf);
            if (m->mbuf) hipFree(m->mbuf);
            delete m;
        }
    for (auto* g : e->grids)
        if (g) {
            if (g->dbuf) hipFree(g->dbuf);
            if (g->mbuf) hipFree(g->mbuf);
            delete g->mom;
            delete g;
        }
    if (e->d_list) hipFree(e->d_list);
    if (e->d_table) hipFree(e->d_table);
    if (e->ev0) hipEventDestroy(e->ev0);
    if (e->ev1) hipEventDestroy(e->ev1);
    if (e->ev_fork) hipEventDestroy(e->ev_fork);
    if (e->ev_join) hipEventDestroy(e->ev_join);
    if (e->s2) hipStreamDestroy(e->s2);
    if (e->s) hipStreamDestroy(e->s);
    delete e;
    return CGO_OK;
}

int cgo_last_error(const cgo_engine* e, char* buf, size_t n) {
    if (!e || !buf || n == 0) return CGO_E_ARG;
    snprintf(buf, n, "%s", e->err.c_str());
    return CGO_OK;
}

}  // extern "C"

// Wait for everything enqueued on the engine stream, then free the deferred
// buffers: no member has work in flight afterwards.
static int drain(cgo_engine* e) {
    HIPCHK(e, hipStreamSynchronize(e->s));
    for (auto* x : e->members)
        if (x) x->queued = false;
    for (void* p : e->graveyard) hipFree(p);
    e->graveyard.clear();
    return CGO_OK;
}

// Before the host reads or rewrites member m's device memory: wait for the
// stream only if m may have work in flight.  A member being set up (created
// since the stream last drained and not yet stepped) is filled while the
// other members step.
static int settle(cgo_engine* e, MemberH* m) { return m->queued ? drain(e) : CGO_OK; }

extern "C" {

int cgo_sync(cgo_engine* e) {
    if (!e) return CGO_E_ARG;
    return drain(e);
}

int cgo_grid_create(cgo_engine* e, const cgo_grid_desc* d, int32_t* grid_id) {
    if (!e || !d || !grid_id || !d->k1) return fail(e, CGO_E_ARG, "grid_create: null argument");
    const int mi = d->maxi, mj = d->maxj, mk = d->maxk;
    if (mi < 3 || mj < 3 || mk < 2 || mk > MAXK_DEV) return fail(e, CGO_E_ARG, "grid_create: bad dims %d %d %d", mi, mj, mk);
    // a flux tile of one column and its neighbours (2 maxi + 1) must fit one
    // LDS slot of a 128-thread block (flux_tw(128) positions)
    if (2 * mi + 1 > flux_tw(128)) return fail(e, CGO_E_ARG, "grid_create: maxi %d above %d", mi, (flux_tw(128) - 1) / 2);
    const double* jarr[] = {d->rc, d->rc2, d->cv, d->cv2, d->rds, d->dsv, d->rdsv, d->ds, d->sv};
    const double* karr[] = {d->dt, d->rdz, d->rdza, d->dza, d->ssmax, d->dz};
    for (auto* p : jarr)
        if (!p) return fail(e, CGO_E_ARG, "grid_create: null per-j metric");
    for (auto* p : karr)
        if (!p) return fail(e, CGO_E_ARG, "grid_create: null per-k metric");
    GridH* g = new GridH();
    g->d = *d;
    const int JS = mj + 2, KS = mk + 2, ncol = mi * mj, ncell = ncol * mk;
    g->k1.assign(d->k1, d->k1 + (size_t)JS * (mi + 2));
    g->mj.assign((size_t)NMJ * JS, 0.0);
    g->mk.assign((size_t)NMK * KS, 0.0);
    for (int j = 0; j < JS; j++) {
        g->mj[MJ_RC * JS + j] = d->rc[j];
        g->mj[MJ_RC2 * JS + j] = d->rc2[j];
        g->mj[MJ_CV * JS + j] = d->cv[j];
        g->mj[MJ_CV2 * JS + j] = d->cv2[j];
        g->mj[MJ_RDS * JS + j] = d->rds[j];
        g->mj[MJ_DSVM * JS + j] = (j >= 1) ? d->dsv[std::min(j, mj - 1)] : 0.0;
        g->mj[MJ_RDSV * JS + j] = d->rdsv[j];
        g->mj[MJ_DS * JS + j] = d->ds[j];
        g->mj[MJ_SV * JS + j] = d->sv[j];
    }
    // BIOGEM geometry, sub_init_phys_ocn (biogem_data.f90:1098-1140)
    std::vector<double> dzl(KS + 1, 0.0), dzal(KS + 1, 0.0);
    for (int k = 1; k <= mk; k++) {
        dzl[k] = d->dz[k];
        dzal[k] = d->dza[k];
    }
    dzal[mk] = dzl[mk] / 2.0;
    for (int k = 0; k < KS; k++) {
        g->mk[MK_DT * KS + k] = d->dt[k];
        g->mk[MK_RDZ * KS + k] = d->rdz[k];
        g->mk[MK_RDZA * KS + k] = d->rdza[k];
        g->mk[MK_DZA * KS + k] = d->dza[k];
        g->mk[MK_SSMAX * KS + k] = d->ssmax[k];
        g->mk[MK_DZ * KS + k] = d->dz[k];
    }
    for (int k = 1; k <= mk; k++) {
        double sm = 0.0, sb = 0.0;
        for (int q = k; q <= mk; q++) sm += d->dsc * dzal[q];
        for (int q = k; q <= mk; q++) sb += d->dsc * dzl[q];
        g->mk[MK_DMID * KS + k] = sm;
        g->mk[MK_DBOT * KS + k] = sb;
        g->mk[MK_DD * KS + k] = d->dsc * dzl[k];
    }
    // wet columns in the reference vector order (i outer, j inner; biogem_lib.f90:1069-1080)
    auto K1 = [&](int I, int J) { return g->k1[(size_t)J * (mi + 2) + I]; };
    for (int I = 1; I <= mi; I++)
        for (int J = 1; J <= mj; J++)
            if (K1(I, J) <= mk) g->wet.push_back((I - 1) + mi * (J - 1));
    {
        std::vector<std::pair<int32_t, int32_t>> o;
        for (size_t q = 0; q < g->wet.size(); q++) o.push_back({g->wet[q], (int32_t)q});
        std::sort(o.begin(), o.end());
        for (auto& pr : o) {
            g->wets.push_back(pr.first);
            g->vpos.push_back(pr.second);
        }
    }
    g->V.assign(ncell, 0.0);
    g->invw.assign(ncell, 0.0);
    for (int I = 1; I <= mi; I++)
        for (int J = 1; J <= mj; J++)
            for (int k = K1(I, J); k <= mk; k++) {
                const double A = 2.0 * CONST_PI * (CONST_REARTH * CONST_REARTH) * (1.0 / mi) * (d->sv[J] - d->sv[J - 1]);
                const size_t c = (size_t)(k - 1) * ncol + (I - 1) + (size_t)mi * (J - 1);
                g->V[c] = g->mk[MK_DD * KS + k] * A;
                g->invw[c] = d->dphi * d->ds[J] * d->dz[k];
            }
    g->Vtot = 0.0;
    for (int32_t c : g->wet) {
        const int I = c % mi + 1, J = c / mi + 1;
        double part = 0.0;
        for (int k = K1(I, J); k <= mk; k++) part += g->V[(size_t)(k - 1) * ncol + c];
        g->Vtot += part;
    }
    // one device buffer: k1 | mj | mk | mjf | mkf | wet | V | invw
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_k1 = take(g->k1.size() * 4), o_mj = take(g->mj.size() * 8), o_mk = take(g->mk.size() * 8);
    const size_t o_mjf = take(g->mj.size() * 4), o_mkf = take(g->mk.size() * 4), o_wet = take(g->wet.size() * 4 + 4);
    const size_t o_wets = take(g->wet.size() * 4 + 4), o_vpos = take(g->wet.size() * 4 + 4);
    const size_t o_V = take((size_t)ncell * 8), o_w = take((size_t)ncell * 8), o_cs = take((size_t)ncol * 4);
    // flux blocks: up to nt consecutive wet columns (ascending), cut early so
    // the staged range with its neighbours stays within flux_tw(nt)
    std::vector<FluxBlk> fb[2];
    for (int v = 0; v < 2; v++) {
        const int nt = v ? 256 : 128, cap = flux_tw(nt);
        const int nw = (int)g->wets.size();
        auto t0of = [&](int w) { return std::max(g->wets[w] - mi, 0); };
        auto t1of = [&](int w) { return std::min(g->wets[w] + mi + 1, ncol); };
        for (int w = 0; w < nw;) {
            int e2 = w + 1;
            while (e2 < nw && e2 - w < nt && t1of(e2) - t0of(w) <= cap) e2++;
            fb[v].push_back(FluxBlk{w, e2, t0of(w), t1of(e2 - 1)});
            g->ftw[v] = std::max(g->ftw[v], t1of(e2 - 1) - t0of(w));
            w = e2;
        }
        // the flux kernel's LDS slots hold cap positions
        if (g->ftw[v] > cap) {
            const int wide = g->ftw[v];
            delete g;
            return fail(e, CGO_E_ARG, "grid_create: a flux block stages %d positions (at most %d)", wide, cap);
        }
    }
    const size_t o_fb0 = take(fb[0].size() * sizeof(FluxBlk) + 16), o_fb1 = take(fb[1].size() * sizeof(FluxBlk) + 16);
    hipSetDevice(e->device);
    if (hipMalloc(&g->dbuf, off) != hipSuccess) {
        delete g;
        return fail(e, CGO_E_NOMEM, "grid_create: hipMalloc(%zu)", off);
    }
    char* b = static_cast<char*>(g->dbuf);
    std::vector<float> mjf(g->mj.begin(), g->mj.end()), mkf(g->mk.begin(), g->mk.end());
    HIPCHK(e, hipMemcpy(b + o_k1, g->k1.data(), g->k1.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(b + o_mj, g->mj.data(), g->mj.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(b + o_mk, g->mk.data(), g->mk.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(b + o_mjf, mjf.data(), mjf.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(b + o_mkf, mkf.data(), mkf.size() * 4, hipMemcpyHostToDevice));
    if (!g->wet.empty()) {
        HIPCHK(e, hipMemcpy(b + o_wet, g->wet.data(), g->wet.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(e, hipMemcpy(b + o_wets, g->wets.data(), g->wet.size() * 4, hipMemcpyHostToDevice));
        HIPCHK(e, hipMemcpy(b + o_vpos, g->vpos.data(), g->wet.size() * 4, hipMemcpyHostToDevice));
    }
    HIPCHK(e, hipMemcpy(b + o_V, g->V.data(), (size_t)ncell * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(b + o_w, g->invw.data(), (size_t)ncell * 8, hipMemcpyHostToDevice));
    if (!fb[0].empty()) HIPCHK(e, hipMemcpy(b + o_fb0, fb[0].data(), fb[0].size() * sizeof(FluxBlk), hipMemcpyHostToDevice));
    if (!fb[1].empty()) HIPCHK(e, hipMemcpy(b + o_fb1, fb[1].data(), fb[1].size() * sizeof(FluxBlk), hipMemcpyHostToDevice));
    {
        std::vector<int32_t> cs(ncol, -1);
        for (size_t q = 0; q < g->wet.size(); q++) cs[g->wet[q]] = (int32_t)q;
        HIPCHK(e, hipMemcpy(b + o_cs, cs.data(), (size_t)ncol * 4, hipMemcpyHostToDevice));
    }
    GridDev& gd = g->dev;
    gd.maxi = mi;
    gd.maxj = mj;
    gd.maxk = mk;
    gd.ncol = ncol;
    gd.ncell = ncell;
    gd.nwet = (int)g->wet.size();
    gd.k1 = reinterpret_cast<const int32_t*>(b + o_k1);
    gd.mj = reinterpret_cast<const double*>(b + o_mj);
    gd.mk = reinterpret_cast<const double*>(b + o_mk);
    gd.mjf = reinterpret_cast<const float*>(b + o_mjf);
    gd.mkf = reinterpret_cast<const float*>(b + o_mkf);
    gd.wetcols = reinterpret_cast<const int32_t*>(b + o_wet);
    gd.wetsorted = reinterpret_cast<const int32_t*>(b + o_wets);
    gd.vpos = reinterpret_cast<const int32_t*>(b + o_vpos);
    gd.V = reinterpret_cast<const double*>(b + o_V);
    gd.cslot = reinterpret_cast<const int32_t*>(b + o_cs);
    gd.fblk[0] = reinterpret_cast<const FluxBlk*>(b + o_fb0);
    gd.fblk[1] = reinterpret_cast<const FluxBlk*>(b + o_fb1);
    gd.nfblk[0] = (int)fb[0].size();
    gd.nfblk[1] = (int)fb[1].size();
    g->d_invw = reinterpret_cast<double*>(b + o_w);
    gd.dphi = d->dphi;
    gd.rdphi = d->rdphi;
    gd.rVtot = g->Vtot > 0 ? 1.0 / g->Vtot : 0.0;
    for (int q = 0; q < 5; q++) gd.ec[q] = d->ec[q];
    {
        // FNV-1a over the dimensions, the mask and the metric tables
        uint64_t h = 1469598103934665603ull;
        auto mix = [&](const void* p, size_t n) {
            const unsigned char* b = static_cast<const unsigned char*>(p);
            for (size_t q = 0; q < n; q++) h = (h ^ b[q]) * 1099511628211ull;
        };
        const int32_t dims[3] = {mi, mj, mk};
        mix(dims, sizeof dims);
        mix(g->k1.data(), g->k1.size() * 4);
        mix(g->mj.data(), g->mj.size() * 8);
        mix(g->mk.data(), g->mk.size() * 8);
        g->fingerprint = h;
    }
    e->grids.push_back(g);
    *grid_id = (int32_t)e->grids.size() - 1;
    return CGO_OK;
}

// byte layout of one member allocation
struct MemLayout {
    size_t st[2], lo[2], u, sflux, rho, cost, ctrl, isoc, ocn, part, settle, carb, M, rM, vd, light, mld, wspeed, atm,
        fatm, colsum, colold, scal, rtab, mink, total;
};

static MemLayout layout(const GridH* g, int L, int prec, int carry, int bgc) {
    MemLayout m{};
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    const size_t ncell = g->dev.ncell, ncol = g->dev.ncol, nwet = g->dev.nwet;
    const size_t es = prec == CGO_FP64 ? 8 : 4;
    m.st[0] = take(L * ncell * es);
    m.st[1] = take(L * ncell * es);
    const bool lo = prec == CGO_FP32C && carry;
    m.lo[0] = lo ? take(L * ncell * 2) : 0;
    m.lo[1] = lo ? take(L * ncell * 2) : 0;
    // FP64: u, v, w; FP32C: the six face coefficients of the flux and the divergence
    m.u = take((prec == CGO_FP64 ? 3 : 7) * ncell * es);
    m.sflux = take(L * ncol * 8);
    m.rho = take(ncell * 8);
    m.cost = take(ncol * 8);
    m.ctrl = take(256);  // parity, steps, diag[4], dmax, sfresh
    m.isoc = take((size_t)10 * ncell * 8);
    if (bgc) {
        m.ocn = take((2 + NBGC) * ncell * 8);
        m.part = take(NSED * ncell * 8);
        m.settle = take(NSED * ncol * 8);
        m.carb = take(NCARB * ncol * 8);
        m.M = take(ncell * 8);
        m.rM = take(ncell * 8);
        m.vd = take(NBGC * ncell * 8);
        m.light = take(ncol * 8);
        m.mld = take(ncol * 8);
        m.wspeed = take(ncol * 8);
        m.atm = take(2 * ncol * 8);
        m.fatm = take(2 * ncol * 8);
        m.colsum = take((L + 1) * nwet * 8);
        m.colold = take((L + 1) * nwet * 8);
        m.scal = take((SC_FACT0 + L + 1) * 8);
        m.rtab = take((size_t)RT_N * (g->dev.maxk + 2) * 8);
        m.mink = take((size_t)(g->dev.maxk + 2) * 4);
    }
    m.total = off;
    return m;
}

int64_t cgo_member_bytes(cgo_engine* e, int32_t grid_id, int32_t maxl, int32_t precision, int32_t bgc) {
    if (!e || grid_id < 0 || grid_id >= (int)e->grids.size()) return CGO_E_ARG;
    return (int64_t)layout(e->grids[grid_id], maxl, precision, 1, bgc).total;
}

static int alloc_member(cgo_engine* e, MemberH* m) {
    const GridH* g = e->grids[m->grid];
    MemLayout ly = layout(g, m->L, m->prec, m->p.carry, m->bgc);
    void* old = m->dbuf;
    const bool had_sflux = m->dev.sflux != nullptr;
    std::vector<char> keep;
    MemLayout oldly = layout(g, m->L, m->prec, m->p.carry, 0);
    if (old) {  // growing to the BIOGEM layout keeps the GOLDSTEIN part
        keep.resize(oldly.total);
        HIPCHK(e, hipMemcpy(keep.data(), old, oldly.total, hipMemcpyDeviceToHost));
    }
    void* buf = nullptr;
    if (hipMalloc(&buf, ly.total) != hipSuccess) return fail(e, CGO_E_NOMEM, "member: hipMalloc(%zu) failed", ly.total);
    HIPCHK(e, hipMemset(buf, 0, ly.total));
    if (old) {
        HIPCHK(e, hipMemcpy(buf, keep.data(), oldly.total, hipMemcpyHostToDevice));
        e->graveyard.push_back(old);  // the caller settled the member: nothing in flight reads it
    }
    m->dbuf = buf;
    m->bytes = ly.total;
    char* b = static_cast<char*>(buf);
    MemberDev& d = m->dev;
    d.g = g->dev;
    d.L = m->L;
    d.prec = m->prec;
    d.diso = m->p.diso;
    d.bgc = m->bgc;
    d.diff1 = m->p.diff[0];
    d.diff2 = m->p.diff[1];
    d.diff1f = (float)m->p.diff[0];
    d.diff2f = (float)m->p.diff[1];
    for (int q = 0; q < 2; q++) {
        d.st[q] = b + ly.st[q];
        d.lo[q] = (m->prec == CGO_FP32C && m->p.carry) ? reinterpret_cast<int16_t*>(b + ly.lo[q]) : nullptr;
    }
    d.u = b + ly.u;
    d.sflux = nullptr;
    d.rho = reinterpret_cast<double*>(b + ly.rho);
    d.cost = reinterpret_cast<double*>(b + ly.cost);
    d.isoc = reinterpret_cast<double*>(b + ly.isoc);
    d.parity = reinterpret_cast<int32_t*>(b + ly.ctrl);
    d.steps = reinterpret_cast<int64_t*>(b + ly.ctrl + 8);
    d.diag = reinterpret_cast<int32_t*>(b + ly.ctrl + 16);
    d.dmax = reinterpret_cast<double*>(b + ly.ctrl + 32);
    d.sfresh = reinterpret_cast<int32_t*>(b + ly.ctrl + 40);
    if (m->bgc) {
        d.ocn = reinterpret_cast<double*>(b + ly.ocn);
        d.part = reinterpret_cast<double*>(b + ly.part);
        d.settle = reinterpret_cast<double*>(b + ly.settle);
        d.carb = reinterpret_cast<double*>(b + ly.carb);
        d.M = reinterpret_cast<double*>(b + ly.M);
        d.rM = reinterpret_cast<double*>(b + ly.rM);
        d.vd = reinterpret_cast<double*>(b + ly.vd);
        d.light = reinterpret_cast<double*>(b + ly.light);
        d.mld = reinterpret_cast<double*>(b + ly.mld);
        d.wspeed = reinterpret_cast<double*>(b + ly.wspeed);
        d.atm = reinterpret_cast<double*>(b + ly.atm);
        d.fatm = reinterpret_cast<double*>(b + ly.fatm);
        d.colsum = reinterpret_cast<double*>(b + ly.colsum);
        d.colold = reinterpret_cast<double*>(b + ly.colold);
        d.scal = reinterpret_cast<double*>(b + ly.scal);
        d.rtab = reinterpret_cast<const double*>(b + ly.rtab);
        d.mink = reinterpret_cast<const int32_t*>(b + ly.mink);
    }
    m->dev.sflux = had_sflux ? reinterpret_cast<double*>(b + ly.sflux) : nullptr;
    d.sfw = m->sbuf ? reinterpret_cast<double*>(b + ly.sflux) : nullptr;
    if (m->mbuf && m->prec == CGO_FP64) d.u64 = reinterpret_cast<double*>(b + ly.u);
    e->table_dirty = true;
    return CGO_OK;
}

int cgo_member_create(cgo_engine* e, int32_t grid_id, int32_t maxl, int32_t precision, const cgo_member_params* p,
                      int32_t* member_id) {
    if (!e || !p || !member_id) return fail(e, CGO_E_ARG, "member_create: null argument");
    if (grid_id < 0 || grid_id >= (int)e->grids.size()) return fail(e, CGO_E_ARG, "member_create: bad grid %d", grid_id);
    if (maxl < 2 || maxl > 256) return fail(e, CGO_E_ARG, "member_create: maxl %d outside 2..256", maxl);
    if (precision != CGO_FP64 && precision != CGO_FP32C) return fail(e, CGO_E_ARG, "member_create: bad precision");
    // kernels address per-member arrays with 32-bit byte offsets
    if ((size_t)maxl * e->grids[grid_id]->dev.ncell * sizeof(double) >= ((size_t)1 << 31))
        return fail(e, CGO_E_ARG, "member_create: %d tracers x %d cells exceed 2 GiB per state array", maxl,
                    e->grids[grid_id]->dev.ncell);
    int slot = -1;
    for (size_t i = 0; i < e->members.size(); i++)
        if (!e->members[i] || !e->members[i]->alive) {
            slot = (int)i;
            break;
        }
    if (slot < 0) {
        if ((int)e->members.size() >= TABLE_CAP) return fail(e, CGO_E_NOMEM, "member_create: table full");
        e->members.push_back(nullptr);
        slot = (int)e->members.size() - 1;
    }
    MemberH* m = e->members[slot] ? e->members[slot] : new MemberH();
    *m = MemberH();
    m->grid = grid_id;
    m->L = maxl;
    m->prec = precision;
    m->p = *p;
    e->members[slot] = m;
    hipSetDevice(e->device);
    int rc = alloc_member(e, m);
    if (rc) return rc;
    m->alive = true;
    e->table_dirty = true;
    bump_gen(e, slot);
    *member_id = slot;
    return CGO_OK;
}

int cgo_member_destroy(cgo_engine* e, int32_t id) {
    MemberH* m = member(e, id);
    if (!m) return fail(e, CGO_E_ARG, "member_destroy: bad member %d", id);
    int rc = settle(e, m);
    if (rc) return rc;
    if (m->dbuf) e->graveyard.push_back(m->dbuf);
    if (m->mbuf) e->graveyard.push_back(m->mbuf);
    if (m->sbuf) e->graveyard.push_back(m->sbuf);
    m->dbuf = nullptr;
    m->mbuf = nullptr;
    m->sbuf = nullptr;
    m->alive = false;
    e->table_dirty = true;
    bump_gen(e, id);
    return CGO_OK;
}

int cgo_set_velocity(cgo_engine* e, int32_t id, const double* u) {
    MemberH* m = member(e, id);
    if (!m || !u) return fail(e, CGO_E_ARG, "set_velocity: bad argument");
    if (int rc = settle(e, m)) return rc;
    if (m->mbuf) {  // prescribed velocities from now on: the member's solve is switched off
        e->graveyard.push_back(m->mbuf);
        m->mbuf = nullptr;
        m->mbytes = 0;
        m->dev.mg = nullptr;
        m->dev.mw = nullptr;
        m->dev.u64 = nullptr;
        e->table_dirty = true;
        bump_gen(e, id);
    }
    const GridDev& g = m->dev.g;
    const int mi = g.maxi, mj = g.maxj, mk = g.maxk;
    const size_t ncell = g.ncell;
    std::vector<double> h(3 * ncell);
    // u(n, 0:maxi, 0:maxj, maxk): element (n,i,j,k) at n + 3*(i + (maxi+1)*(j + (maxj+1)*(k-1)))
    for (int k = 1; k <= mk; k++)
        for (int J = 1; J <= mj; J++)
            for (int I = 1; I <= mi; I++)
                for (int n = 0; n < 3; n++)
                    h[(size_t)n * ncell + (size_t)(k - 1) * g.ncol + (I - 1) + (size_t)mi * (J - 1)] =
                        u[n + 3 * ((size_t)I + (size_t)(mi + 1) * ((size_t)J + (size_t)(mj + 1) * (k - 1)))];
    if (m->prec == CGO_FP64) {
        HIPCHK(e, hipMemcpy(const_cast<void*>(m->dev.u), h.data(), h.size() * 8, hipMemcpyHostToDevice));
    } else {
        // FP32 form: per wet cell the six face coefficients of the flux in
        // coefficient form (the face flux is the neighbour difference times
        // the coefficient) with the Peclet-weighted upwinding of
        // goldstein.f90:2489-2523 folded in, and the discrete divergence of
        // the open faces, all evaluated in FP64 from the FP64 velocities and
        // rounded once (the flux kernel works with q - q_cell and adds
        // q_cell * div u)
        const GridH* gh = e->grids[m->grid];
        const int JS = mj + 2, KS = mk + 2;
        const size_t ncol = g.ncol;
        auto K1 = [&](int I, int J) { return gh->k1[(size_t)J * (mi + 2) + I]; };
        const double diff1 = m->p.diff[0], diffv = m->p.diff[1];
        auto upwind = [](double pec) { return pec / (2.0 + std::fabs(pec)); };
        std::vector<float> f(7 * ncell, 0.0f);
        for (int k = 1; k <= mk; k++)
            for (int J = 1; J <= mj; J++)
                for (int I = 1; I <= mi; I++) {
                    const int k1c = K1(I, J);
                    if (k < k1c) continue;
                    const size_t o = (size_t)(k - 1) * ncol;
                    const size_t c = o + (I - 1) + (size_t)mi * (J - 1);
                    const size_t cw = o + (size_t)mi * (J - 1) + (I == 1 ? mi - 1 : I - 2);
                    const size_t cs = o + (size_t)mi * (J > 1 ? J - 2 : 0) + (I - 1);
                    const bool oe = k >= std::max(k1c, K1(I + 1, J)), ow = k >= std::max(K1(I - 1, J), k1c);
                    const bool on = k >= std::max(k1c, K1(I, J + 1)), os = k >= std::max(K1(I, J - 1), k1c);
                    const double rc = gh->mj[MJ_RC * JS + J], rds = gh->mj[MJ_RDS * JS + J];
                    const double cvn = gh->mj[MJ_CV * JS + J], cvs = gh->mj[MJ_CV * JS + J - 1];
                    const double rdz = gh->mk[MK_RDZ * KS + k];
                    const double ue = h[c], uw = h[cw], un = h[ncell + c], us = h[ncell + cs];
                    const double ua = k == mk ? 0.0 : h[2 * ncell + c];
                    const double wb = k > k1c ? h[2 * ncell + c - ncol] : 0.0;
                    const double div = ((oe ? ue * rc : 0.0) - (ow ? uw * rc : 0.0)) * g.rdphi +
                                       ((on ? cvn * un : 0.0) - (os ? cvs * us : 0.0)) * rds + (ua - wb) * rdz;
                    const double rc2 = gh->mj[MJ_RC2 * JS + J];
                    const double cv2n = gh->mj[MJ_CV2 * JS + J], cv2s = gh->mj[MJ_CV2 * JS + J - 1];
                    const double dsvn = gh->mj[MJ_DSVM * JS + J], dsvs = gh->mj[MJ_DSVM * JS + J - 1];
                    const double dza = gh->mk[MK_DZA * KS + k], rdza = gh->mk[MK_RDZA * KS + k];
                    const double upe = upwind(ue * g.dphi / diff1), upw = upwind(uw * g.dphi / diff1);
                    const double upn = upwind(un * dsvn / diff1), ups = upwind(us * dsvs / diff1);
                    const double upa = upwind(ua * dza / diffv);
                    const bool top = k == mk;
                    f[c] = oe ? (float)((ue * rc * (1.0 - upe) * 0.5 - rc2 * diff1) * g.rdphi) : 0.0f;
                    f[ncell + c] = ow ? (float)((uw * rc * (1.0 + upw) * 0.5 + rc2 * diff1) * g.rdphi) : 0.0f;
                    f[2 * ncell + c] = on ? (float)((cvn * un * (1.0 - upn) * 0.5 - cv2n * diff1) * rds) : 0.0f;
                    f[3 * ncell + c] = os ? (float)((cvs * us * (1.0 + ups) * 0.5 + cv2s * diff1) * rds) : 0.0f;
                    f[4 * ncell + c] = top ? 0.0f : (float)(ua * (1.0 - upa) * 0.5 - rdza * diffv);
                    f[5 * ncell + c] = top ? 0.0f : (float)(-(ua * (1.0 + upa) * 0.5) - rdza * diffv);
                    f[6 * ncell + c] = (float)div;
                }
        HIPCHK(e, hipMemcpy(const_cast<void*>(m->dev.u), f.data(), f.size() * 4, hipMemcpyHostToDevice));
    }
    return CGO_OK;
}

// ------------------------------------------------------------ velocity solve

static constexpr double MOM_FSC = 2 * 7.2921E-5;  // fsc, goldstein_lib.f90:49
static constexpr double MOM_USC = 0.05;           // usc
static constexpr double MOM_RH0SC = 1.0E3;        // rh0sc, goldstein_lib.f90:50

int cgo_grid_set_momentum(cgo_engine* e, int32_t grid_id, const cgo_momentum_desc* d) {
    if (!e || !d || grid_id < 0 || grid_id >= (int)e->grids.size()) return fail(e, CGO_E_ARG, "grid_set_momentum: bad argument");
    if (!d->s || !d->c || !d->rcv || !d->rds2 || !d->labels || (d->isles > 0 && (!d->npi || !d->paths)))
        return fail(e, CGO_E_ARG, "grid_set_momentum: null array");
    GridH* g = e->grids[grid_id];
    if (g->mom) return fail(e, CGO_E_STATE, "grid_set_momentum: grid %d already has its velocity solve", grid_id);
    const int mi = g->d.maxi, mj = g->d.maxj, mk = g->d.maxk, JS = mj + 2, KS = mk + 2;
    // the band solve runs on one wave: its 64 lanes hold the pivot row and
    // the maxi+1 rows below it that a step of the elimination updates
    if (mi > 62) return fail(e, CGO_E_ARG, "grid_set_momentum: maxi %d above 62", mi);
    // the solve block's LDS: the psi forcing, the staged products and the
    // ring of factor rows
    if (mom_solve_lds(mi * (mj + 1), mi) > 65536)
        return fail(e, CGO_E_ARG, "grid_set_momentum: %d psi points with maxi %d need %zu bytes of LDS, above 65536",
                    mi * (mj + 1), mi, mom_solve_lds(mi * (mj + 1), mi));
    if (!(d->adrag > 0.0) || !(d->rel >= 0.0 && d->rel <= 1.0))
        return fail(e, CGO_E_ARG, "grid_set_momentum: adrag %g days, rel %g", d->adrag, d->rel);
    const double* jrows[JT_N];
    jrows[JT_S] = d->s;
    jrows[JT_SV] = &g->mj[MJ_SV * JS];
    jrows[JT_C] = d->c;
    jrows[JT_CV] = &g->mj[MJ_CV * JS];
    jrows[JT_RC] = &g->mj[MJ_RC * JS];
    jrows[JT_RCV] = d->rcv;
    jrows[JT_RDS] = &g->mj[MJ_RDS * JS];
    jrows[JT_RDSV] = &g->mj[MJ_RDSV * JS];
    jrows[JT_RDS2] = d->rds2;
    jrows[JT_DS] = &g->mj[MJ_DS * JS];
    jrows[JT_DSV] = &g->mj[MJ_DSVM * JS];  // dsv(j) at j = 1..maxj-1, the only rows the solve reads it at
    MomSetup* M = new MomSetup();
    const std::string err = mom_setup(*M, g->k1.data(), mi, mj, mk, jrows, &g->mk[MK_DZ * KS], &g->mk[MK_DZA * KS],
                                      g->d.dphi, g->d.rdphi, d->labels, d->isles, d->npi, d->paths, d->adrag, MOM_FSC);
    if (!err.empty()) {
        delete M;
        return fail(e, CGO_E_ARG, "grid_set_momentum: %s", err.c_str());
    }
    const int n = M->n, N = M->N, nis = std::max(M->isles, 1);
    {
        size_t np = 0;
        for (int q = 0; q < M->isles; q++) np += (size_t)M->npi[q];
        if (2 * np > mom_ring_terms(n)) {
            delete M;
            return fail(e, CGO_E_ARG, "grid_set_momentum: %zu island path points, at most %zu", np, mom_ring_terms(n) / 2);
        }
    }
    // the elimination's multipliers by pivot row and the eliminated band by row
    std::vector<double> fwd((size_t)std::max(N - 1, 1) * (n + 1), 0.0), bwd((size_t)N * (n + 2), 0.0);
    for (int i = 1; i <= N - 1; i++)
        for (int q = 1; q <= n + 1 && i + q <= N; q++) fwd[(size_t)(i - 1) * (n + 1) + q - 1] = M->RATM(i + q, q);
    std::vector<uint64_t> bmask(N, 0);
    for (int i = 1; i <= N; i++)
        for (int k = 0; k <= n + 1 && i + k <= N; k++) {
            bwd[(size_t)(i - 1) * (n + 2) + k] = M->GAP(i, n + 2 + k);
            if (k >= 1 && M->GAP(i, n + 2 + k) != 0.0) bmask[i - 1] |= (uint64_t)1 << (k - 1);
        }
    // one device buffer: the MomGridDev struct, then its arrays
    size_t off = 0;
    auto take = [&](size_t bytes) {
        size_t o = off;
        off += (bytes + 255) & ~(size_t)255;
        return o;
    };
    struct Piece {
        const void* src;
        size_t bytes, off;
    };
    std::vector<Piece> pieces;
    auto add = [&](const void* src, size_t bytes) {
        pieces.push_back({src, bytes, take(bytes)});
        return pieces.size() - 1;
    };
    const size_t o_struct = take(sizeof(MomGridDev));
    const size_t p_jt = add(M->jt.data(), M->jt.size() * 8), p_dz = add(M->dz.data(), M->dz.size() * 8),
                 p_dza = add(M->dza.data(), M->dza.size() * 8), p_rh = add(M->rh.data(), M->rh.size() * 8),
                 p_drag = add(M->drag.data(), M->drag.size() * 8), p_rtv = add(M->rtv.data(), M->rtv.size() * 8),
                 p_rtv3 = add(M->rtv3.data(), M->rtv3.size() * 8), p_ku = add(M->ku.data(), M->ku.size() * 4),
                 p_mk = add(M->mkv.data(), M->mkv.size() * 4), p_getj = add(M->getj.data(), M->getj.size()),
                 p_fwd = add(fwd.data(), fwd.size() * 8), p_bwd = add(bwd.data(), bwd.size() * 8),
                 p_bm = add(bmask.data(), bmask.size() * 8),
                 p_ubisl = add(M->ubisl.data(), M->ubisl.size() * 8), p_psisl = add(M->psisl.data(), M->psisl.size() * 8),
                 p_erisl = add(M->erisl.data(), M->erisl.size() * 8), p_npi = add(M->npi.data(), (size_t)nis * 4),
                 p_lp = add(M->lpisl.data(), M->lpisl.size() * 4), p_ip = add(M->ipisl.data(), M->ipisl.size() * 4),
                 p_jp = add(M->jpisl.data(), M->jpisl.size() * 4);
    // the psi equation's inverse for the FP32 form's dense solve, where it is
    // small enough to keep (N <= MOM_DENSE_MAXN: at most 128 MiB)
    std::vector<double> ainv;
    size_t p_ainv = SIZE_MAX;
    if (N <= MOM_DENSE_MAXN) {
        mom_inverse(*M, ainv);
        p_ainv = add(ainv.data(), ainv.size() * 8);
    }
    hipSetDevice(e->device);
    void* buf = nullptr;
    if (hipMalloc(&buf, off) != hipSuccess) {
        delete M;
        return fail(e, CGO_E_NOMEM, "grid_set_momentum: hipMalloc(%zu)", off);
    }
    char* b = static_cast<char*>(buf);
    auto at_ = [&](size_t p) { return b + pieces[p].off; };
    MomGridDev h{};
    h.maxi = mi;
    h.maxj = mj;
    h.maxk = mk;
    h.n = n;
    h.N = N;
    h.isles = M->isles;
    h.mpi = M->mpi;
    h.dphi = g->d.dphi;
    h.rdphi = g->d.rdphi;
    h.rel = d->rel;
    h.jt = (const double*)at_(p_jt);
    h.dz = (const double*)at_(p_dz);
    h.dza = (const double*)at_(p_dza);
    h.rh = (const double*)at_(p_rh);
    h.drag = (const double*)at_(p_drag);
    h.rtv = (const double*)at_(p_rtv);
    h.rtv3 = (const double*)at_(p_rtv3);
    h.ku = (const int32_t*)at_(p_ku);
    h.mk = (const int32_t*)at_(p_mk);
    h.getj = (const uint8_t*)at_(p_getj);
    h.fwd = (const double*)at_(p_fwd);
    h.bwd = (const double*)at_(p_bwd);
    h.bmask = (const uint64_t*)at_(p_bm);
    h.ubisl = (const double*)at_(p_ubisl);
    h.psisl = (const double*)at_(p_psisl);
    h.erisl = (const double*)at_(p_erisl);
    h.npi = (const int32_t*)at_(p_npi);
    h.lpisl = (const int32_t*)at_(p_lp);
    h.ipisl = (const int32_t*)at_(p_ip);
    h.jpisl = (const int32_t*)at_(p_jp);
    h.ainv = p_ainv == SIZE_MAX ? nullptr : (const double*)at_(p_ainv);
    int rc = CGO_OK;
    if (hipMemcpy(b + o_struct, &h, sizeof h, hipMemcpyHostToDevice) != hipSuccess) rc = CGO_E_HIP;
    for (const Piece& pc : pieces)
        if (rc == CGO_OK && pc.bytes && hipMemcpy(b + pc.off, pc.src, pc.bytes, hipMemcpyHostToDevice) != hipSuccess)
            rc = CGO_E_HIP;
    if (rc != CGO_OK) {
        hipFree(buf);
        delete M;
        return fail(e, CGO_E_HIP, "grid_set_momentum: upload failed");
    }
    // fingerprint of the set-up's inputs: a checkpoint of a member with the
    // solve restores only onto a grid set up alike
    uint64_t fp = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t nb) {
        const unsigned char* q = static_cast<const unsigned char*>(p);
        for (size_t x = 0; x < nb; x++) fp = (fp ^ q[x]) * 1099511628211ull;
    };
    mix(M->jt.data(), M->jt.size() * 8);
    mix(d->labels, (size_t)mi * (mj + 1) * 4);
    mix(&d->isles, 4);
    mix(M->npi.data(), (size_t)nis * 4);
    mix(M->lpisl.data(), M->lpisl.size() * 4);
    mix(M->ipisl.data(), M->ipisl.size() * 4);
    mix(M->jpisl.data(), M->jpisl.size() * 4);
    mix(&d->adrag, 8);
    mix(&d->rel, 8);
    g->mom = M;
    g->mbuf = buf;
    g->d_mg = reinterpret_cast<const MomGridDev*>(b + o_struct);
    g->mom_fp = fp;
    g->ainv_on = p_ainv != SIZE_MAX;
    return CGO_OK;
}

// the member's arrays of the solve, allocated and zeroed, and its u planes
// zeroed: the solve leaves closed faces, and w at the top level, at zero
static int mom_attach(cgo_engine* e, MemberH* m, int32_t id) {
    const GridH* g = e->grids[m->grid];
    const MomSetup& S = *g->mom;
    const MomMemberLayout L = mom_member_layout(S.mi, S.mj, S.mk, m->prec != CGO_FP64);
    void* buf = nullptr;
    if (hipMalloc(&buf, L.total * 8) != hipSuccess) return fail(e, CGO_E_NOMEM, "set_wind: hipMalloc(%zu)", L.total * 8);
    if (hipMemset(buf, 0, L.total * 8) != hipSuccess ||
        hipMemset(const_cast<void*>(m->dev.u), 0, (size_t)m->dev.g.ncell * (m->prec == CGO_FP64 ? 3 * 8 : 7 * 4)) != hipSuccess) {
        hipFree(buf);
        return fail(e, CGO_E_HIP, "set_wind: hipMemset failed");
    }
    m->mbuf = buf;
    m->mbytes = L.total * 8;
    m->dev.mg = g->d_mg;
    m->dev.mw = static_cast<double*>(buf);
    m->dev.u64 = m->prec == CGO_FP64 ? static_cast<double*>(const_cast<void*>(m->dev.u)) : static_cast<double*>(buf) + L.u64;
    e->table_dirty = true;
    bump_gen(e, id);
    return CGO_OK;
}

// ---------------------------------------------------------------- surface fluxes
static int download_dev(cgo_engine* e, MemberH* m, std::vector<double>& h);
// embm_lib.f90:37-85: the scales of the wind stress and the bulk drag
static constexpr double SFX_DSC = 5.0E3, SFX_RHOAIR = 1.25, SFX_CD = 0.0013;

// usurf, the wind speed of the bulk formulas, from the last wind stress
// (surflux, embm.f90:2756-2810): the stress's magnitude at the T point from
// its x at the u points either side and y at the v points either side,
// inverted through |tau| = rho_air cd u^2; the two rows nearest each pole
// take their zonal mean unless par_wind_polar_avg is 2
static void sfx_usurf(const GridH* g, const std::vector<double>& st, std::vector<double>& us) {
    const int mi = g->d.maxi, mj = g->d.maxj, mk = g->d.maxk, KS = mk + 2;
    const size_t ncol = (size_t)mi * mj;
    us.assign(ncol, 0.0);
    if (st.size() != 4 * ncol) return;
    const double dzz = g->mk[MK_DZ * KS + mk] * g->mk[MK_DZA * KS + mk - 1] / 2;
    const double den = MOM_RH0SC * SFX_DSC * MOM_USC * MOM_FSC;
    const double scf = g->sfx_scf;
    std::vector<double> t1(ncol), t2(ncol);
    for (size_t c = 0; c < ncol; c++) {
        t1[c] = scf * st[c] / den / dzz * dzz;
        t2[c] = scf * st[3 * ncol + c] / den / dzz * dzz;
    }
    for (int j = 1; j <= mj; j++) {
        double tv3 = 0.0;
        for (int i = 1; i <= mi; i++) {
            const size_t c = (size_t)(i - 1) + (size_t)mi * (j - 1);
            const size_t cw = i == 1 ? (size_t)(mi - 1) + (size_t)mi * (j - 1) : c - 1;
            const double tv = (t1[c] + t1[cw]) / 2;
            const double tv2 = j == 1 ? t2[c] / 2 : (t2[c] + t2[c - mi]) / 2;
            us[c] = sqrt(sqrt(tv * tv + tv2 * tv2) * MOM_RH0SC * SFX_DSC * MOM_USC * MOM_FSC / (SFX_RHOAIR * SFX_CD * scf));
            tv3 = tv3 + us[c];
        }
        if (g->sfx_polar != 2 && (j <= 2 || j >= mj - 1))
            for (int i = 1; i <= mi; i++) us[(size_t)(i - 1) + (size_t)mi * (j - 1)] = tv3 / mi;
    }
}

int cgo_grid_set_surflux(cgo_engine* e, int32_t grid_id, const cgo_surflux_desc* d) {
    if (!e || !d || grid_id < 0 || grid_id >= (int)e->grids.size()) return fail(e, CGO_E_ARG, "grid_set_surflux: bad argument");
    if (!d->solfor || !d->albcl || !d->ca || !d->iroff || !d->jroff || d->nyear < 1)
        return fail(e, CGO_E_ARG, "grid_set_surflux: null table or nyear %d", d->nyear);
    GridH* g = e->grids[grid_id];
    if (g->sbuf) return fail(e, CGO_E_STATE, "grid_set_surflux: grid %d already has its surface fluxes", grid_id);
    const int mi = g->d.maxi, mj = g->d.maxj, mk = g->d.maxk;
    const int ncol = mi * mj;
    if (sfx_lds(ncol) > 65536) return fail(e, CGO_E_ARG, "grid_set_surflux: %d columns above the 8192 one workgroup holds", ncol);
    auto K1 = [&](int I, int J) { return g->k1[(size_t)J * (mi + 2) + I]; };
    // runoff gather lists: each wet column's land sources in the order the
    // reference's loop visits them (i outer, j inner, embm.f90:2955-2956)
    std::vector<std::vector<int32_t>> src(ncol);
    for (int i = 1; i <= mi; i++)
        for (int j = 1; j <= mj; j++) {
            if (K1(i, j) <= mk) continue;
            const size_t c = (size_t)(i - 1) + (size_t)mi * (j - 1);
            const int ir = d->iroff[c], jr = d->jroff[c];
            if (ir < 1 || ir > mi || jr < 1 || jr > mj || K1(ir, jr) > mk)
                return fail(e, CGO_E_ARG, "grid_set_surflux: runoff of (%d,%d) routed to (%d,%d), not an ocean cell", i, j, ir, jr);
            src[(size_t)(ir - 1) + (size_t)mi * (jr - 1)].push_back((int32_t)c);
        }
    std::vector<int32_t> rstart(ncol + 1, 0), rsrc;
    for (int c = 0; c < ncol; c++) {
        rstart[c] = (int32_t)rsrc.size();
        rsrc.insert(rsrc.end(), src[c].begin(), src[c].end());
    }
    rstart[ncol] = (int32_t)rsrc.size();
    std::vector<double> pm(ncol, 0.0);
    if (d->pmeadj) std::copy(d->pmeadj, d->pmeadj + ncol, pm.begin());
    // one allocation: the struct, then the tables
    size_t off = 0;
    auto take = [&](size_t nb) {
        size_t o = off;
        off += (nb + 255) & ~(size_t)255;
        return o;
    };
    const size_t o_s = take(sizeof(SfxGridDev)), o_sol = take((size_t)d->nyear * mj * 8), o_alb = take((size_t)ncol * 8),
                 o_ca = take((size_t)ncol * 8), o_pm = take((size_t)ncol * 8), o_rs = take((size_t)(ncol + 1) * 4),
                 o_src = take(std::max<size_t>(rsrc.size(), 1) * 4);
    void* buf = nullptr;
    hipSetDevice(e->device);
    if (hipMalloc(&buf, off) != hipSuccess) return fail(e, CGO_E_NOMEM, "grid_set_surflux: hipMalloc(%zu)", off);
    char* b = static_cast<char*>(buf);
    SfxGridDev h{};
    h.nyear = d->nyear;
    h.igrid = d->igrid;
    h.radfor_atchem = d->radfor_atchem ? 1 : 0;
    h.nsrc = (int32_t)rsrc.size();
    h.rmax = d->rmax;
    h.delf2x = d->delf2x;
    h.rdtdim = d->rdtdim;
    h.rsictscsf = d->rsictscsf;
    h.olr_adj = d->olr_adj;
    h.olr_adj0 = d->olr_adj0;
    h.t_eqm = d->t_eqm;
    h.albsic_min = d->albsic_min;
    h.albsic_max = d->albsic_max;
    h.sich_max = d->sich_max;
    h.rate_co2 = d->rate_co2;
    h.rate_ch4 = d->rate_ch4;
    h.rate_n2o = d->rate_n2o;
    h.solfor = reinterpret_cast<const double*>(b + o_sol);
    h.albcl = reinterpret_cast<const double*>(b + o_alb);
    h.ca = reinterpret_cast<const double*>(b + o_ca);
    h.pmeadj = reinterpret_cast<const double*>(b + o_pm);
    h.rstart = reinterpret_cast<const int32_t*>(b + o_rs);
    h.rsrc = reinterpret_cast<const int32_t*>(b + o_src);
    bool ok = hipMemcpy(b + o_s, &h, sizeof h, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(b + o_sol, d->solfor, (size_t)d->nyear * mj * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(b + o_alb, d->albcl, (size_t)ncol * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(b + o_ca, d->ca, (size_t)ncol * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(b + o_pm, pm.data(), (size_t)ncol * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(b + o_rs, rstart.data(), rstart.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
              (rsrc.empty() || hipMemcpy(b + o_src, rsrc.data(), rsrc.size() * 4, hipMemcpyHostToDevice) == hipSuccess);
    if (!ok) {
        hipFree(buf);
        return fail(e, CGO_E_HIP, "grid_set_surflux: upload failed");
    }
    uint64_t fp = 1469598103934665603ull;
    auto mix = [&](const void* p, size_t nb) {
        const unsigned char* q = static_cast<const unsigned char*>(p);
        for (size_t x = 0; x < nb; x++) fp = (fp ^ q[x]) * 1099511628211ull;
    };
    mix(&h.nyear, 4 * sizeof(int32_t));
    mix(&h.rmax, 13 * sizeof(double));
    mix(d->solfor, (size_t)d->nyear * mj * 8);
    mix(d->albcl, (size_t)ncol * 8);
    mix(d->ca, (size_t)ncol * 8);
    mix(pm.data(), (size_t)ncol * 8);
    mix(rsrc.data(), rsrc.size() * 4);
    mix(&d->scf, 8);
    mix(&d->polar_avg, 4);
    g->sbuf = buf;
    g->d_sg = reinterpret_cast<const SfxGridDev*>(b + o_s);
    g->sfx_scf = d->scf;
    g->sfx_polar = d->polar_avg;
    g->sfx_atchem = h.radfor_atchem;
    g->sfx_gas[0] = d->co2;
    g->sfx_gas[1] = d->ch4;
    g->sfx_gas[2] = d->n2o;
    g->sfx_fp = fp;
    return CGO_OK;
}

// the top level's T and S of a state in device order [l][k][j][i]
static void sfx_top(const GridDev& g, const std::vector<double>& st, double* t, double* s) {
    const size_t top = (size_t)(g.maxk - 1) * g.ncol;
    for (int c = 0; c < g.ncol; c++) {
        t[c] = st[top + c];
        s[c] = st[(size_t)g.ncell + top + c];
    }
}

int cgo_set_surface(cgo_engine* e, int32_t id, const double* atemp, const double* ashum, const double* sich,
                    const double* sica, const double* tice, const double* fx0sic, const double* fwsic) {
    MemberH* m = member(e, id);
    if (!m) return fail(e, CGO_E_ARG, "set_surface: bad member %d", id);
    const GridH* g = e->grids[m->grid];
    if (int rc = settle(e, m)) return rc;
    hipSetDevice(e->device);
    if (!atemp) {
        if (m->sbuf) {
            e->graveyard.push_back(m->sbuf);
            m->sbuf = nullptr;
            m->sbytes = 0;
            m->dev.sg = nullptr;
            m->dev.sw = nullptr;
            m->dev.sfw = nullptr;
            m->dev.sflux = nullptr;
            e->table_dirty = true;
            bump_gen(e, id);
        }
        return CGO_OK;
    }
    if (!g->sbuf) return fail(e, CGO_E_STATE, "set_surface: grid %d has no surface fluxes (cgo_grid_set_surflux)", m->grid);
    if (!ashum) return fail(e, CGO_E_ARG, "set_surface: ashum is needed with atemp");
    const size_t ncol = (size_t)m->dev.g.ncol;
    const size_t nd = sfx_member_doubles((int)ncol);
    std::vector<double> h(SW_SSTAR * ncol + ncol, 0.0);
    const double* src[SW_USURF] = {atemp, ashum, sich, sica, tice};
    for (int f = 0; f < SW_USURF; f++)
        if (src[f]) std::copy(src[f], src[f] + ncol, h.begin() + f * ncol);
    if (fx0sic) std::copy(fx0sic, fx0sic + ncol, h.begin() + SW_FX0SIC * ncol);
    if (fwsic) std::copy(fwsic, fwsic + ncol, h.begin() + SW_FWSIC * ncol);
    std::vector<double> us;
    sfx_usurf(g, m->stress, us);
    std::copy(us.begin(), us.end(), h.begin() + SW_USURF * ncol);
    if (m->sbuf) {
        // already on: new air and sea-ice fields; the top level, the outputs,
        // the gases and the call count stay
        HIPCHK(e, hipMemcpy(m->sbuf, h.data(), SW_TSTAR * ncol * 8, hipMemcpyHostToDevice));
        return CGO_OK;
    }
    std::vector<double> st;
    if (int rc = download_dev(e, m, st)) return rc;
    sfx_top(m->dev.g, st, &h[SW_TSTAR * ncol], &h[SW_SSTAR * ncol]);
    std::vector<double> tail(nd - h.size(), 0.0);
    double* sc = tail.data() + SF_N * ncol;
    sc[SS_CO2] = g->sfx_gas[0];
    sc[SS_CH4] = g->sfx_gas[1];
    sc[SS_N2O] = g->sfx_gas[2];
    sc[SS_ISTEP] = 0.0;
    void* buf = nullptr;
    if (hipMalloc(&buf, nd * 8) != hipSuccess) return fail(e, CGO_E_NOMEM, "set_surface: hipMalloc(%zu)", nd * 8);
    const MemLayout ly = layout(g, m->L, m->prec, m->p.carry, m->bgc);
    double* slot = reinterpret_cast<double*>(static_cast<char*>(m->dbuf) + ly.sflux);
    if (hipMemcpy(buf, h.data(), h.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(static_cast<double*>(buf) + h.size(), tail.data(), tail.size() * 8, hipMemcpyHostToDevice) != hipSuccess ||
        hipMemset(slot, 0, (size_t)m->L * ncol * 8) != hipSuccess) {
        hipFree(buf);
        return fail(e, CGO_E_HIP, "set_surface: upload failed");
    }
    m->sbuf = buf;
    m->sbytes = nd * 8;
    m->dev.sg = g->d_sg;
    m->dev.sw = static_cast<double*>(buf);
    m->dev.sfw = slot;
    m->dev.sflux = slot;
    e->table_dirty = true;
    bump_gen(e, id);
    return CGO_OK;
}

int cgo_download_surflux(cgo_engine* e, int32_t id, double* out, double* ashum, double* tice, double* gases) {
    MemberH* m = member(e, id);
    if (!m) return fail(e, CGO_E_ARG, "download_surflux: bad member %d", id);
    if (!m->sbuf) return fail(e, CGO_E_STATE, "download_surflux: member %d has no surface fluxes (cgo_set_surface)", id);
    if (int rc = settle(e, m)) return rc;
    const size_t ncol = (size_t)m->dev.g.ncol;
    const double* w = static_cast<const double*>(m->sbuf);
    if (out) HIPCHK(e, hipMemcpy(out, w + SW_NF * ncol, SF_N * ncol * 8, hipMemcpyDeviceToHost));
    if (ashum) HIPCHK(e, hipMemcpy(ashum, w + SW_ASHUM * ncol, ncol * 8, hipMemcpyDeviceToHost));
    if (tice) HIPCHK(e, hipMemcpy(tice, w + SW_TICE * ncol, ncol * 8, hipMemcpyDeviceToHost));
    if (gases) HIPCHK(e, hipMemcpy(gases, w + (SW_NF + SF_N) * ncol, SS_N * 8, hipMemcpyDeviceToHost));
    return CGO_OK;
}

int cgo_set_wind(cgo_engine* e, int32_t id, const double* sxu, const double* syu, const double* sxv, const double* syv,
                 double scf) {
    MemberH* m = member(e, id);
    if (!m || !sxu || !syu || !sxv || !syv) return fail(e, CGO_E_ARG, "set_wind: bad argument");
    const GridH* g = e->grids[m->grid];
    if (!g->mom && !g->sbuf)
        return fail(e, CGO_E_STATE, "set_wind: grid %d has neither a velocity solve nor surface fluxes", m->grid);
    if (int rc = settle(e, m)) return rc;
    hipSetDevice(e->device);
    {
        // kept for the surface fluxes' wind speed (also when they are switched on later)
        const size_t nc = (size_t)m->dev.g.ncol;
        m->stress.resize(4 * nc);
        const double* f[4] = {sxu, syu, sxv, syv};
        for (int q = 0; q < 4; q++) std::copy(f[q], f[q] + nc, m->stress.begin() + q * nc);
        if (m->sbuf) {
            std::vector<double> us;
            sfx_usurf(g, m->stress, us);
            HIPCHK(e, hipMemcpy(static_cast<double*>(m->sbuf) + SW_USURF * nc, us.data(), nc * 8, hipMemcpyHostToDevice));
        }
    }
    if (!g->mom) return CGO_OK;
    if (!m->mbuf)
        if (int rc = mom_attach(e, m, id)) return rc;
    const MomSetup& S = *g->mom;
    const int mi = S.mi, mj = S.mj, mk = S.mk, n = S.n;
    const size_t ncol = (size_t)mi * mj;
    const MomMemberLayout L = mom_member_layout(mi, mj, mk, m->prec != CGO_FP64);
    // dztau, dztav: the stresses scaled by scf and made non-dimensional per
    // unit depth of the top layer's half-cell dzz; tau their x at u and y at
    // v points (step_goldstein 102-107, dzz 1035)
    const double dzz = S.dz[mk] * S.dza[mk - 1] / 2;
    const double den = MOM_RH0SC * g->d.dsc * MOM_USC * MOM_FSC;
    std::vector<double> tau(2 * ncol), dztau(2 * ncol), dztav(2 * ncol), gbw(S.N, 0.0);
    for (size_t c = 0; c < ncol; c++) {
        dztau[2 * c] = scf * sxu[c] / den / dzz;
        dztau[2 * c + 1] = scf * syu[c] / den / dzz;
        dztav[2 * c] = scf * sxv[c] / den / dzz;
        dztav[2 * c + 1] = scf * syv[c] / den / dzz;
        tau[2 * c] = dztau[2 * c] * dzz;
        tau[2 * c + 1] = dztav[2 * c + 1] * dzz;
    }
    // the wind's curl at every wet psi point (wind, 3685-3714)
    auto T = [&](int l, int i, int j) { return tau[(l - 1) + 2 * ((size_t)(i - 1) + (size_t)mi * (j - 1))]; };
    for (int i = 1; i <= mi; i++)
        for (int j = 0; j <= mj; j++) {
            const int ip1 = i % mi + 1;
            if (std::max(std::max(S.K1(i, j), S.K1(i + 1, j)), std::max(S.K1(i, j + 1), S.K1(i + 1, j + 1))) <= mk)
                gbw[(i + j * n) - 1] = (T(2, ip1, j) * S.RH(2, i + 1, j) - T(2, i, j) * S.RH(2, i, j)) * S.rdphi * S.J(JT_RCV, j) -
                                       (T(1, i, j + 1) * S.J(JT_C, j + 1) * S.RH(1, i, j + 1) -
                                        T(1, i, j) * S.J(JT_C, j) * S.RH(1, i, j)) *
                                           S.J(JT_RDSV, j);
        }
    double* w = m->dev.mw;
    HIPCHK(e, hipMemcpy(w + L.gbw, gbw.data(), gbw.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(w + L.tau, tau.data(), tau.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(w + L.dztau, dztau.data(), dztau.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(w + L.dztav, dztav.data(), dztav.size() * 8, hipMemcpyHostToDevice));
    return CGO_OK;
}

int cgo_download_velocity(cgo_engine* e, int32_t id, double* go_u) {
    MemberH* m = member(e, id);
    if (!m || !go_u) return fail(e, CGO_E_ARG, "download_velocity: bad argument");
    if (m->prec != CGO_FP64 && !m->mbuf)
        return fail(e, CGO_E_STATE, "download_velocity: member %d holds its prescribed velocities as FP32 flux coefficients", id);
    if (int rc = settle(e, m)) return rc;
    const GridDev& g = m->dev.g;
    const size_t ncell = g.ncell;
    std::vector<double> h(3 * ncell);
    const void* src = m->prec == CGO_FP64 ? m->dev.u : m->dev.u64;
    HIPCHK(e, hipMemcpy(h.data(), src, h.size() * 8, hipMemcpyDeviceToHost));
    // [l][k][j][i] -> go_u(3, maxi, maxj, maxk)
    for (size_t c = 0; c < ncell; c++)
        for (int l = 0; l < 3; l++) go_u[(size_t)l + 3 * c] = h[(size_t)l * ncell + c];
    return CGO_OK;
}

int cgo_download_psi(cgo_engine* e, int32_t id, double* go_psi) {
    MemberH* m = member(e, id);
    if (!m || !go_psi) return fail(e, CGO_E_ARG, "download_psi: bad argument");
    if (!m->mbuf) return fail(e, CGO_E_STATE, "download_psi: member %d has no velocity solve (cgo_set_wind)", id);
    if (int rc = settle(e, m)) return rc;
    const GridH* g = e->grids[m->grid];
    const MomMemberLayout L = mom_member_layout(g->mom->mi, g->mom->mj, g->mom->mk, m->prec != CGO_FP64);
    HIPCHK(e, hipMemcpy(go_psi, m->dev.mw + L.psi, (size_t)(g->mom->mi + 1) * (g->mom->mj + 1) * 8, hipMemcpyDeviceToHost));
    return CGO_OK;
}

int cgo_set_sflux(cgo_engine* e, int32_t id, const double* sflux) {
    MemberH* m = member(e, id);
    if (!m) return fail(e, CGO_E_ARG, "set_sflux: bad member");
    if (m->sbuf)
        return fail(e, CGO_E_STATE, "set_sflux: member %d computes its surface slot (cgo_set_surface); switch that off first", id);
    if (int rc = settle(e, m)) return rc;
    const GridH* g = e->grids[m->grid];
    MemLayout ly = layout(g, m->L, m->prec, m->p.carry, m->bgc);
    if (!sflux) {
        m->dev.sflux = nullptr;
    } else {
        const int mi = g->dev.maxi, mj = g->dev.maxj, L = m->L;
        std::vector<double> h((size_t)L * g->dev.ncol);
        for (int J = 0; J < mj; J++)
            for (int I = 0; I < mi; I++)
                for (int l = 0; l < L; l++)
                    h[(size_t)l * g->dev.ncol + I + (size_t)mi * J] = sflux[l + (size_t)L * (I + (size_t)mi * J)];
        double* dst = reinterpret_cast<double*>(static_cast<char*>(m->dbuf) + ly.sflux);
        HIPCHK(e, hipMemcpy(dst, h.data(), h.size() * 8, hipMemcpyHostToDevice));
        m->dev.sflux = dst;
    }
    e->table_dirty = true;
    return CGO_OK;
}

// go_ts(maxl,maxi,maxj,maxk) <-> device [l][k][j][i]
static void to_device_order(const GridDev& g, int L, const double* go, std::vector<double>& h) {
    const size_t ncell = g.ncell;
    h.resize((size_t)L * ncell);
    for (size_t c = 0; c < ncell; c++)
        for (int l = 0; l < L; l++) h[(size_t)l * ncell + c] = go[(size_t)l + (size_t)L * c];
}

int cgo_upload_state(cgo_engine* e, int32_t id, const double* go_ts) {
    MemberH* m = member(e, id);
    if (!m || !go_ts) return fail(e, CGO_E_ARG, "upload_state: bad argument");
    if (int rc = settle(e, m)) return rc;
    const GridDev& g = m->dev.g;
    std::vector<double> h;
    to_device_order(g, m->L, go_ts, h);
    const size_t n = h.size();
    std::vector<float> f;
    std::vector<int16_t> lo;
    if (m->prec != CGO_FP64) {
        f.resize(n);
        lo.resize(n);
        for (size_t i = 0; i < n; i++) {
            f[i] = (float)h[i];
            // (h - f) / carry_unit(f), the division by a power of two done exactly
            const int ex = f[i] == 0.0f ? -126 : std::max(ilogbf(f[i]), -126);
            const double r = rint(ldexp(h[i] - (double)f[i], 39 - ex));
            lo[i] = (int16_t)std::min(std::max(r, -32767.0), 32767.0);
        }
    }
    for (int q = 0; q < 2; q++) {
        if (m->prec == CGO_FP64) {
            HIPCHK(e, hipMemcpy(m->dev.st[q], h.data(), n * 8, hipMemcpyHostToDevice));
        } else {
            HIPCHK(e, hipMemcpy(m->dev.st[q], f.data(), n * 4, hipMemcpyHostToDevice));
            if (m->dev.lo[q]) HIPCHK(e, hipMemcpy(m->dev.lo[q], lo.data(), n * 2, hipMemcpyHostToDevice));
        }
    }
    // rho of the new state at every wet cell, as initialise_goldstein forms
    // it from the initial or restart state (goldstein.f90:1450, 1754), for
    // the velocity solve's first step.  Once the member has stepped, rho is
    // the last tstepo's, which jbar reads whatever the host does to go_ts in
    // between (step_goldstein's ts = go_ts, 176-177, leaves rho alone).
    if (!m->stepped) {
        const GridH* gh = e->grids[m->grid];
        const int mi = g.maxi, mj = g.maxj, mk = g.maxk;
        std::vector<double> r(g.ncell, 0.0);
        for (int k = 1; k <= mk; k++)
            for (int J = 1; J <= mj; J++)
                for (int I = 1; I <= mi; I++)
                    if (k >= gh->k1[(size_t)J * (mi + 2) + I]) {
                        const size_t c = (size_t)(k - 1) * g.ncol + (I - 1) + (size_t)mi * (J - 1);
                        const double t = h[c], sa = h[g.ncell + c];
                        r[c] = g.ec[0] * t + g.ec[1] * sa + g.ec[2] * (t * t) + g.ec[3] * (t * t * t);
                    }
        HIPCHK(e, hipMemcpy(m->dev.rho, r.data(), r.size() * 8, hipMemcpyHostToDevice));
    }
    // the surface fluxes read the top level of the state uploaded (as
    // initialise_goldstein hands tstar_ocn over, goldstein.f90:2015-2017)
    // until the next tstepo replaces it
    if (m->sbuf) {
        std::vector<double> top(2 * (size_t)g.ncol);
        sfx_top(g, h, top.data(), top.data() + g.ncol);
        HIPCHK(e, hipMemcpy(static_cast<double*>(m->sbuf) + SW_TSTAR * (size_t)g.ncol, top.data(), top.size() * 8,
                            hipMemcpyHostToDevice));
    }
    const int32_t zero = 0;
    HIPCHK(e, hipMemcpy(m->dev.parity, &zero, 4, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(m->dev.sfresh, &zero, 4, hipMemcpyHostToDevice));
    return CGO_OK;
}

static int download_dev(cgo_engine* e, MemberH* m, std::vector<double>& h) {
    const GridDev& g = m->dev.g;
    const size_t n = (size_t)m->L * g.ncell;
    int32_t par = 0;
    if (int rc = settle(e, m)) return rc;
    HIPCHK(e, hipMemcpy(&par, m->dev.parity, 4, hipMemcpyDeviceToHost));
    h.resize(n);
    if (m->prec == CGO_FP64) {
        HIPCHK(e, hipMemcpy(h.data(), m->dev.st[par], n * 8, hipMemcpyDeviceToHost));
    } else {
        std::vector<float> f(n);
        std::vector<int16_t> lo(n, 0);
        HIPCHK(e, hipMemcpy(f.data(), m->dev.st[par], n * 4, hipMemcpyDeviceToHost));
        if (m->dev.lo[par]) HIPCHK(e, hipMemcpy(lo.data(), m->dev.lo[par], n * 2, hipMemcpyDeviceToHost));
        for (size_t i = 0; i < n; i++) h[i] = (double)f[i] + (double)lo[i] * carry_unit_h(f[i]);
    }
    return CGO_OK;
}

int cgo_download_state(cgo_engine* e, int32_t id, double* go_ts) {
    MemberH* m = member(e, id);
    if (!m || !go_ts) return fail(e, CGO_E_ARG, "download_state: bad argument");
    std::vector<double> h;
    int rc = download_dev(e, m, h);
    if (rc) return rc;
    const size_t ncell = m->dev.g.ncell;
    for (size_t c = 0; c < ncell; c++)
        for (int l = 0; l < m->L; l++) go_ts[(size_t)l + (size_t)m->L * c] = h[(size_t)l * ncell + c];
    return CGO_OK;
}

int cgo_download_rho(cgo_engine* e, int32_t id, double* rho) {
    MemberH* m = member(e, id);
    if (!m || !rho) return fail(e, CGO_E_ARG, "download_rho: bad argument");
    HIPCHK(e, hipStreamSynchronize(e->s));
    HIPCHK(e, hipMemcpy(rho, m->dev.rho, (size_t)m->dev.g.ncell * 8, hipMemcpyDeviceToHost));
    return CGO_OK;
}

int cgo_inventory(cgo_engine* e, int32_t id, double* inv) {
    MemberH* m = member(e, id);
    if (!m || !inv) return fail(e, CGO_E_ARG, "inventory: bad argument");
    int rc = sync_table(e);
    if (rc) return rc;
    double* d_out = nullptr;
    HIPCHK(e, hipMalloc(&d_out, (size_t)m->L * 8));
    launch_inventory(e->d_table, id, m->L, d_out, e->grids[m->grid]->d_invw, m->prec, e->s);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipStreamSynchronize(e->s));
    HIPCHK(e, hipMemcpy(inv, d_out, (size_t)m->L * 8, hipMemcpyDeviceToHost));
    hipFree(d_out);
    return CGO_OK;
}

int cgo_inventory_noalloc(cgo_engine* e, int32_t id, double* inv) {
    MemberH* m = member(e, id);
    if (!m || !inv) return fail(e, CGO_E_ARG, "inventory_noalloc: bad argument");
    if (m->L > NOALLOC_MAXL)
        return fail(e, CGO_E_ARG, "inventory_noalloc: member %d has %d tracers, at most %d without allocation", id, m->L,
                    NOALLOC_MAXL);
    double* ptr = e->device < NOALLOC_MAXDEV ? noalloc_ptr[e->device] : nullptr;
    if (!ptr) return fail(e, CGO_E_STATE, "inventory_noalloc: no scratch address for device %d", e->device);
    HIPCHK(e, hipSetDevice(e->device));
    if (e->table_dirty) {
        // bring this member's table entry up to date from the host copy (the
        // whole-table refresh would build a host vector); work in flight
        // reads the table, so it finishes first
        HIPCHK(e, hipStreamSynchronize(e->s));
        HIPCHK(e, hipMemcpy(e->d_table + id, &m->dev, sizeof(MemberDev), hipMemcpyHostToDevice));
    }
    // every one of the member's L sums is written by the launch: nothing to clear
    launch_inventory(e->d_table, id, m->L, ptr, e->grids[m->grid]->d_invw, m->prec, e->s);
    HIPCHK(e, hipGetLastError());
    HIPCHK(e, hipStreamSynchronize(e->s));
    HIPCHK(e, hipMemcpy(inv, ptr, (size_t)m->L * 8, hipMemcpyDeviceToHost));
    return CGO_OK;
}

// the device copy of the BIOGEM parameters, with each carried species' slot
// in the state and its plane in the ocn block
static void set_bgc_params(MemberH* m, const cgo_bgc_params& bp) {
    m->bp = bp;
    BgcParams& P = m->dev.bp;
    const int32_t slots[NBGC] = {bp.slot_DIC, bp.slot_PO4, bp.slot_O2, bp.slot_ALK, bp.slot_DOM_C, bp.slot_DOM_P};
    int nplane = 2;
    for (int b = 0; b < NBGC; b++) {
        P.slot[b] = slots[b];
        P.ocn[b] = slots[b] >= 0 ? nplane++ : -1;
    }
    P.caco3 = bp.caco3;
    P.dtyr = bp.dtyr;
    P.k0_PO4 = bp.k0_PO4;
    P.c0_PO4 = bp.c0_PO4;
    P.DOMfrac = bp.DOMfrac;
    P.DOMlifetime = bp.DOMlifetime;
    P.red_POP_POC = bp.red_POP_POC;
    P.red_POP_PO2 = bp.red_POP_PO2;
    P.red_PON_ALK = bp.red_PON_ALK;
    P.red_POP_PON = bp.red_POP_PON;
    P.red_POC_CaCO3 = bp.red_POC_CaCO3;
    P.red_POC_CaCO3_pP = bp.red_POC_CaCO3_pP;
    P.POC_frac2 = bp.POC_frac2;
    P.POC_dfrac2 = bp.POC_dfrac2;
    P.POC_c0frac2 = bp.POC_c0frac2;
    P.POC_eL1 = bp.POC_eL1;
    P.POC_eL2 = bp.POC_eL2;
    P.CaCO3_frac2 = bp.CaCO3_frac2;
    P.CaCO3_eL1 = bp.CaCO3_eL1;
    P.CaCO3_eL2 = bp.CaCO3_eL2;
    P.sinkrate = bp.sinkrate;
    P.remin_k_O2 = bp.remin_k_O2;
    P.remin_c0_O2 = bp.remin_c0_O2;
    P.carb_tol = bp.carb_tol;
    P.carb_maxit = bp.carb_maxit;
    P.dsc = bp.dsc;
    P.airsea = bp.airsea;
    P.gas_a = bp.gastransfer_a;
}

// ---------------------------------------------------------- checkpoints
// blob = CkptHdr, then the member's whole device allocation as laid out by
// layout() (state ping-pong and carry planes, velocities, surface flux,
// rho, cost, control words, BIOGEM arrays and scalars), then, for a member
// with a velocity solve, its arrays of the solve (wind, relaxation memory,
// streamfunction): restoring it into a member of the same grid, tracer
// count, precision and carry continues the run bit for bit.
struct CkptHdr {
    char magic[8];
    uint32_t version, hdr_bytes;
    int32_t maxi, maxj, maxk, L, prec, carry, bgc, diso, wstream, coupling_due, has_sflux, mom, sfx, pad;
    uint64_t grid_fp, payload, mom_fp, mom_bytes, sfx_fp, sfx_bytes;
    double diff[2];
    cgo_bgc_params bp;
};
static const char CKPT_MAGIC[8] = {'C', 'G', 'O', 'C', 'K', 'P', 'T', '1'};

int64_t cgo_checkpoint_bytes(cgo_engine* e, int32_t id) {
    MemberH* m = member(e, id);
    if (!m) return fail(e, CGO_E_ARG, "checkpoint_bytes: bad member %d", id);
    return (int64_t)(sizeof(CkptHdr) + m->bytes + m->mbytes + m->sbytes);
}

int cgo_checkpoint(cgo_engine* e, int32_t id, void* buf, size_t n) {
    MemberH* m = member(e, id);
    if (!m || !buf) return fail(e, CGO_E_ARG, "checkpoint: bad argument");
    const size_t need = sizeof(CkptHdr) + m->bytes + m->mbytes + m->sbytes;
    if (n < need) return fail(e, CGO_E_ARG, "checkpoint: buffer of %zu bytes, %zu needed", n, need);
    HIPCHK(e, hipSetDevice(e->device));
    HIPCHK(e, hipStreamSynchronize(e->s));
    CkptHdr h;
    memset(&h, 0, sizeof h);
    memcpy(h.magic, CKPT_MAGIC, 8);
    h.version = 3;
    h.hdr_bytes = (uint32_t)sizeof(CkptHdr);
    const GridDev& g = m->dev.g;
    h.maxi = g.maxi;
    h.maxj = g.maxj;
    h.maxk = g.maxk;
    h.L = m->L;
    h.prec = m->prec;
    h.carry = m->p.carry;
    h.bgc = m->bgc;
    h.diso = m->p.diso;
    h.wstream = m->wstream;
    h.coupling_due = m->coupling_due ? 1 : 0;
    h.has_sflux = m->dev.sflux != nullptr;
    h.grid_fp = e->grids[m->grid]->fingerprint;
    h.payload = m->bytes;
    h.diff[0] = m->p.diff[0];
    h.diff[1] = m->p.diff[1];
    h.bp = m->bp;
    h.mom = m->mbuf != nullptr;
    h.mom_fp = m->mbuf ? e->grids[m->grid]->mom_fp : 0;
    h.mom_bytes = m->mbytes;
    h.sfx = m->sbuf != nullptr;
    h.sfx_fp = m->sbuf ? e->grids[m->grid]->sfx_fp : 0;
    h.sfx_bytes = m->sbytes;
    memcpy(buf, &h, sizeof h);
    HIPCHK(e, hipMemcpy(static_cast<char*>(buf) + sizeof h, m->dbuf, m->bytes, hipMemcpyDeviceToHost));
    if (m->mbuf)
        HIPCHK(e, hipMemcpy(static_cast<char*>(buf) + sizeof h + m->bytes, m->mbuf, m->mbytes, hipMemcpyDeviceToHost));
    if (m->sbuf)
        HIPCHK(e, hipMemcpy(static_cast<char*>(buf) + sizeof h + m->bytes + m->mbytes, m->sbuf, m->sbytes,
                            hipMemcpyDeviceToHost));
    return CGO_OK;
}

int cgo_restore(cgo_engine* e, int32_t id, const void* buf, size_t n) {
    MemberH* m = member(e, id);
    if (!m || !buf || n < sizeof(CkptHdr)) return fail(e, CGO_E_ARG, "restore: bad argument");
    CkptHdr h;
    memcpy(&h, buf, sizeof h);
    if (memcmp(h.magic, CKPT_MAGIC, 8) != 0 || h.version != 3 || h.hdr_bytes != sizeof(CkptHdr))
        return fail(e, CGO_E_ARG, "restore: not a checkpoint of this library version");
    const GridH* gh = e->grids[m->grid];
    if (h.grid_fp != gh->fingerprint || h.L != m->L || h.prec != m->prec || h.carry != m->p.carry)
        return fail(e, CGO_E_ARG, "restore: checkpoint of another grid, tracer count, precision or carry "
                                  "(grid %dx%dx%d, %d tracers, precision %d)", h.maxi, h.maxj, h.maxk, h.L, h.prec);
    if (n < sizeof h + h.payload + h.mom_bytes + h.sfx_bytes)
        return fail(e, CGO_E_ARG, "restore: %zu bytes, the checkpoint holds %zu", n,
                    (size_t)(sizeof h + h.payload + h.mom_bytes + h.sfx_bytes));
    if (h.mom && (!gh->mom || gh->mom_fp != h.mom_fp))
        return fail(e, CGO_E_ARG, "restore: the checkpoint's velocity solve needs its grid set up as it was (cgo_grid_set_momentum)");
    if (h.sfx && (!gh->sbuf || gh->sfx_fp != h.sfx_fp || h.sfx_bytes != sfx_member_doubles(gh->dev.ncol) * 8))
        return fail(e, CGO_E_ARG, "restore: the checkpoint's surface fluxes need its grid set up as it was (cgo_grid_set_surflux)");
    HIPCHK(e, hipSetDevice(e->device));
    if (int rc = settle(e, m)) return rc;
    int32_t slot = id;
    if (h.mom && !m->mbuf) {
        if (int rc = mom_attach(e, m, id)) return rc;
    } else if (!h.mom && m->mbuf) {
        e->graveyard.push_back(m->mbuf);
        m->mbuf = nullptr;
        m->mbytes = 0;
        m->dev.mg = nullptr;
        m->dev.mw = nullptr;
        m->dev.u64 = nullptr;
    }
    if (h.mom && h.mom_bytes != m->mbytes)
        return fail(e, CGO_E_STATE, "restore: the velocity solve's arrays (%zu bytes) differ from the checkpoint's (%zu)",
                    m->mbytes, (size_t)h.mom_bytes);
    if (h.bgc && !m->bgc) {
        m->bgc = 1;
        int rc = alloc_member(e, m);
        if (rc) return rc;
    }
    if (h.payload != m->bytes)
        return fail(e, CGO_E_STATE, "restore: the member's layout (%zu bytes) differs from the checkpoint's (%zu)",
                    m->bytes, (size_t)h.payload);
    HIPCHK(e, hipMemcpy(m->dbuf, static_cast<const char*>(buf) + sizeof h, h.payload, hipMemcpyHostToDevice));
    if (h.mom)
        HIPCHK(e, hipMemcpy(m->mbuf, static_cast<const char*>(buf) + sizeof h + h.payload, h.mom_bytes, hipMemcpyHostToDevice));
    if (h.sfx && !m->sbuf) {
        void* sb = nullptr;
        if (hipMalloc(&sb, h.sfx_bytes) != hipSuccess) return fail(e, CGO_E_NOMEM, "restore: hipMalloc(%zu)", (size_t)h.sfx_bytes);
        m->sbuf = sb;
        m->sbytes = h.sfx_bytes;
    } else if (!h.sfx && m->sbuf) {
        e->graveyard.push_back(m->sbuf);
        m->sbuf = nullptr;
        m->sbytes = 0;
    }
    if (h.sfx)
        HIPCHK(e, hipMemcpy(m->sbuf, static_cast<const char*>(buf) + sizeof h + h.payload + h.mom_bytes, h.sfx_bytes,
                            hipMemcpyHostToDevice));
    m->p.diff[0] = h.diff[0];
    m->p.diff[1] = h.diff[1];
    m->p.diso = h.diso;
    m->dev.diff1 = h.diff[0];
    m->dev.diff2 = h.diff[1];
    m->dev.diff1f = (float)h.diff[0];
    m->dev.diff2f = (float)h.diff[1];
    m->dev.diso = h.diso;
    if (h.bgc) {
        set_bgc_params(m, h.bp);
        m->wstream = h.wstream;
    }
    m->coupling_due = h.coupling_due != 0;
    const MemLayout ly = layout(gh, m->L, m->prec, m->p.carry, m->bgc);
    m->dev.sflux = h.has_sflux ? reinterpret_cast<double*>(static_cast<char*>(m->dbuf) + ly.sflux) : nullptr;
    m->dev.sg = m->sbuf ? gh->d_sg : nullptr;
    m->dev.sw = static_cast<double*>(m->sbuf);
    m->dev.sfw = m->sbuf ? reinterpret_cast<double*>(static_cast<char*>(m->dbuf) + ly.sflux) : nullptr;
    e->table_dirty = true;
    bump_gen(e, slot);
    return CGO_OK;
}

int cgo_set_cost(cgo_engine* e, int32_t id, const double* cost) {
    MemberH* m = member(e, id);
    if (!m || !cost) return fail(e, CGO_E_ARG, "set_cost: bad argument");
    HIPCHK(e, hipStreamSynchronize(e->s));
    HIPCHK(e, hipMemcpy(m->dev.cost, cost, (size_t)m->dev.g.ncol * 8, hipMemcpyHostToDevice));
    return CGO_OK;
}

int cgo_download_cost(cgo_engine* e, int32_t id, double* cost) {
    MemberH* m = member(e, id);
    if (!m || !cost) return fail(e, CGO_E_ARG, "download_cost: bad argument");
    HIPCHK(e, hipStreamSynchronize(e->s));
    HIPCHK(e, hipMemcpy(cost, m->dev.cost, (size_t)m->dev.g.ncol * 8, hipMemcpyDeviceToHost));
    return CGO_OK;
}

int cgo_get_diag(cgo_engine* e, int32_t id, cgo_diag* d) {
    MemberH* m = member(e, id);
    if (!m || !d) return fail(e, CGO_E_ARG, "get_diag: bad argument");
    HIPCHK(e, hipStreamSynchronize(e->s));
    int64_t steps = 0;
    int32_t diag[4] = {0, 0, 0, 0};
    double dmax = 0.0;
    HIPCHK(e, hipMemcpy(&steps, m->dev.steps, 8, hipMemcpyDeviceToHost));
    HIPCHK(e, hipMemcpy(diag, m->dev.diag, 16, hipMemcpyDeviceToHost));
    HIPCHK(e, hipMemcpy(&dmax, m->dev.dmax, 8, hipMemcpyDeviceToHost));
    std::vector<double> cost(m->dev.g.ncol);
    HIPCHK(e, hipMemcpy(cost.data(), m->dev.cost, cost.size() * 8, hipMemcpyDeviceToHost));
    d->steps = steps;
    d->limps = diag[0];
    d->carb_fail = diag[1];
    d->carb_iters = (uint32_t)diag[2];
    d->dmax = dmax;
    d->cost = 0.0;
    for (double c : cost) d->cost += c;
    d->smean = 0.0;
    if (m->bgc) HIPCHK(e, hipMemcpy(&d->smean, m->dev.scal + SC_SMEAN, 8, hipMemcpyDeviceToHost));
    return CGO_OK;
}

void cgo_bgc_default_params(cgo_bgc_params* p, double dtyr) {
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->slot_DIC = p->slot_ALK = p->slot_PO4 = p->slot_O2 = p->slot_DOM_C = p->slot_DOM_P = -1;
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
    p->sinkrate = 125.0 / (1.0 / 365.25);
    p->remin_k_O2 = 1.0;
    p->remin_c0_O2 = 8.0E-6;
    p->carb_tol = 0.001;
    p->carb_maxit = 100;
    p->dsc = 5.0E3;
    p->airsea = 0;
    p->gastransfer_a = 0.310;  // biogem-defaults.nml:221
    p->atm_pCO2 = 278.0E-6;    // atchem-defaults.nml:4
    p->atm_pO2 = 0.2095;
}

int cgo_bgc_init(cgo_engine* e, int32_t id, const cgo_bgc_params* bp, const double* light, const double* mld) {
    MemberH* m = member(e, id);
    if (!m || !bp || !light || !mld) return fail(e, CGO_E_ARG, "bgc_init: bad argument");
    const int32_t slots[NBGC] = {bp->slot_DIC, bp->slot_PO4, bp->slot_O2, bp->slot_ALK, bp->slot_DOM_C, bp->slot_DOM_P};
    for (int b = 0; b < NBGC; b++)
        if (slots[b] >= m->L || slots[b] == 0 || slots[b] == 1 || slots[b] < -1)
            return fail(e, CGO_E_ARG, "bgc_init: species slot %d invalid for maxl %d", slots[b], m->L);
    if (bp->caco3 && (bp->slot_DIC < 0 || bp->slot_ALK < 0)) return fail(e, CGO_E_ARG, "bgc_init: CaCO3 needs DIC and ALK");
    if (m->L > 256) return fail(e, CGO_E_ARG, "bgc_init: at most 256 tracers with BIOGEM (maxl %d)", m->L);
    if (int rc0 = settle(e, m)) return rc0;
    std::vector<double> ts;
    int rc = download_dev(e, m, ts);  // current state, device order
    if (rc) return rc;
    if (!m->bgc) {
        m->bgc = 1;
        rc = alloc_member(e, m);
        bump_gen(e, id);  // new layout, whatever follows
        if (rc) return rc;
    }
    set_bgc_params(m, *bp);
    const BgcParams& P = m->dev.bp;
    // host-side initial state, as sub_init_phys_ocn / sub_init_carb and the
    // inverse of the ts normalisation (biogem.f90:2046)
    const GridH* g = e->grids[m->grid];
    const GridDev& gd = g->dev;
    const int mi = gd.maxi, mk = gd.maxk, L = m->L;
    const size_t ncell = gd.ncell, ncol = gd.ncol, nwet = gd.nwet;
    auto K1 = [&](int I, int J) { return g->k1[(size_t)J * (mi + 2) + I]; };
    std::vector<double> Mh(ncell, 0.0), rMh(ncell, 0.0), Sabs(ncell, 0.0);
    for (int32_t c : g->wet) {
        const int I = c % mi + 1, J = c / mi + 1;
        for (int k = K1(I, J); k <= mk; k++) {
            const size_t cc = (size_t)(k - 1) * ncol + c;
            Mh[cc] = CONV_M3_KG * g->V[cc];
            rMh[cc] = 1.0 / Mh[cc];
            Sabs[cc] = ts[ncell + cc] + SALN0;
        }
    }
    double Smean = 0.0;
    const double rV = gd.rVtot;
    for (int32_t c : g->wet) {
        const int I = c % mi + 1, J = c / mi + 1;
        double part = 0.0;
        for (int k = K1(I, J); k <= mk; k++) part += Sabs[(size_t)(k - 1) * ncol + c] * g->V[(size_t)(k - 1) * ncol + c];
        Smean += part * rV;
    }
    // ocn of every tracer slot (needed for the OLD inventories); planes kept for T, S, species
    std::vector<double> ocn_all((size_t)L * ncell, 0.0);
    for (int32_t c : g->wet) {
        const int I = c % mi + 1, J = c / mi + 1;
        for (int k = K1(I, J); k <= mk; k++) {
            const size_t cc = (size_t)(k - 1) * ncol + c;
            ocn_all[cc] = ts[cc] + ZEROC;
            ocn_all[ncell + cc] = Sabs[cc];
            for (int l = 2; l < L; l++) ocn_all[(size_t)l * ncell + cc] = ts[(size_t)l * ncell + cc] * Sabs[cc] / Smean;
        }
    }
    std::vector<double> planes((size_t)(2 + NBGC) * ncell, 0.0);
    std::copy(ocn_all.begin(), ocn_all.begin() + 2 * ncell, planes.begin());
    for (int b = 0; b < NBGC; b++)
        if (P.slot[b] >= 0)
            std::copy(ocn_all.begin() + (size_t)P.slot[b] * ncell, ocn_all.begin() + (size_t)(P.slot[b] + 1) * ncell,
                      planes.begin() + (size_t)P.ocn[b] * ncell);
    std::vector<double> colold((size_t)(L + 1) * nwet, 0.0);
    for (size_t w = 0; w < nwet; w++) {
        const int c = g->wet[w];
        const int I = c % mi + 1, J = c / mi + 1;
        for (int l = 2; l < L; l++) {
            double s = 0.0;
            for (int k = K1(I, J); k <= mk; k++) {
                const size_t cc = (size_t)(k - 1) * ncol + c;
                s += ocn_all[(size_t)l * ncell + cc] * Mh[cc];
            }
            colold[(size_t)l * nwet + w] = s;
        }
    }
    std::vector<double> carb((size_t)NCARB * ncol, 0.0);
    for (size_t c = 0; c < ncol; c++) carb[IC_H * ncol + c] = pow(10.0, -7.8);
    std::vector<double> lt(ncol), ml(ncol), scal(SC_FACT0 + L + 1, 0.0);
    for (size_t c = 0; c < ncol; c++) {
        lt[c] = light[c];
        ml[c] = mld[c];
    }
    scal[SC_SMEAN] = Smean;
    scal[SC_PSCALE] = 1.0;
    // remineralisation tables, evaluated as sub_box_remin_part does per level
    const int KS = mk + 2;
    const double* Dbot = &g->mk[MK_DBOT * KS];
    const double* dD = &g->mk[MK_DD * KS];
    std::vector<double> rt((size_t)RT_N * KS, 0.0);
    std::vector<int32_t> mink(KS, 0);
    for (int kk = 1; kk <= mk; kk++) {
        if (kk < mk) rt[RT_LAYER * KS + kk] = dD[kk + 1] / dD[kk];
        rt[RT_POC1 * KS + kk] = (1.0 - exp(-dD[kk] / bp->POC_eL1));
        rt[RT_POC2 * KS + kk] = (1.0 - exp(-dD[kk] / bp->POC_eL2));
        rt[RT_CAL1 * KS + kk] = (1.0 - exp(-dD[kk] / bp->CaCO3_eL1));
        rt[RT_CAL2 * KS + kk] = (1.0 - exp(-dD[kk] / bp->CaCO3_eL2));
    }
    for (int k = 1; k <= mk; k++) {
        const double maxD = Dbot[k] + bp->dtyr * bp->sinkrate;
        for (int kk = k - 1; kk >= 1; kk--)
            if (Dbot[kk] > maxD) {
                mink[k] = kk;
                break;
            }
    }
    // Window of the register-resident particle streams: at any level the
    // live streams come from at most W-1 consecutive source levels above it.
    // With multi-level sources (sinking shorter than the ocean depth) a stream
    // from k settles at mink[k], or reaches the floor (mink 0): it spans
    // k - mink[k] levels.  A single surface source needs no window.
    {
        int span = 1;
        if (bp->dtyr * bp->sinkrate <= bp->dsc)
            for (int k = 2; k <= mk; k++) span = std::max(span, k - mink[k]);
        // the smallest compiled window (k_bgc_col<R, W>) that holds the span
        int W = 16;
        for (int w : {4, 5, 6, 8, 16})
            if (w >= span + 1) {
                W = w;
                break;
            }
        if (span + 1 > W)
            return fail(e, CGO_E_ARG, "bgc_init: particles sink through %d levels in one step (at most 15 supported)", span);
        m->wstream = W;
    }
    MemberDev& d = m->dev;
    HIPCHK(e, hipMemcpy(const_cast<double*>(d.rtab), rt.data(), rt.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(const_cast<int32_t*>(d.mink), mink.data(), mink.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(d.ocn, planes.data(), planes.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemset(d.part, 0, (size_t)NSED * ncell * 8));
    HIPCHK(e, hipMemset(d.settle, 0, (size_t)NSED * ncol * 8));
    HIPCHK(e, hipMemset(d.vd, 0, (size_t)NBGC * ncell * 8));
    HIPCHK(e, hipMemcpy(d.carb, carb.data(), carb.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(d.M, Mh.data(), ncell * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(d.rM, rMh.data(), ncell * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(d.light, lt.data(), ncol * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(d.mld, ml.data(), ncol * 8, hipMemcpyHostToDevice));
    // the atmosphere starts uniform (sub_init_tracer_atm_comp, atchem_data.f90:230-260);
    // the wind speed stays as cgo_set_windspeed last set it (zero until then)
    {
        std::vector<double> a(2 * ncol);
        std::fill(a.begin(), a.begin() + ncol, bp->atm_pCO2);
        std::fill(a.begin() + ncol, a.end(), bp->atm_pO2);
        HIPCHK(e, hipMemcpy(d.atm, a.data(), a.size() * 8, hipMemcpyHostToDevice));
        HIPCHK(e, hipMemset(d.fatm, 0, 2 * ncol * 8));
    }
    HIPCHK(e, hipMemcpy(d.colold, colold.data(), colold.size() * 8, hipMemcpyHostToDevice));
    HIPCHK(e, hipMemcpy(d.scal, scal.data(), scal.size() * 8, hipMemcpyHostToDevice));
    const int32_t zero = 0;
    HIPCHK(e, hipMemcpy(d.sfresh, &zero, 4, hipMemcpyHostToDevice));
    e->table_dirty = true;
    bump_gen(e, id);
    return CGO_OK;
}

int cgo_bgc_download(cgo_engine* e, int32_t id, int32_t what, double* out) {
    MemberH* m = member(e, id);
    if (!m || !out || !m->bgc) return fail(e, CGO_E_ARG, "bgc_download: bad argument or no BIOGEM state");
    HIPCHK(e, hipStreamSynchronize(e->s));
    const GridDev& g = m->dev.g;
    const size_t ncell = g.ncell, ncol = g.ncol;
    const int L = m->L;
    const BgcParams& P = m->dev.bp;
    if (what == 0 || what == 3) {
        const bool anom = what == 3;
        const int np = anom ? NBGC : 2 + NBGC;
        std::vector<double> h((size_t)np * ncell);
        HIPCHK(e, hipMemcpy(h.data(), anom ? m->dev.vd : m->dev.ocn, h.size() * 8, hipMemcpyDeviceToHost));
        std::fill(out, out + (size_t)L * ncell, 0.0);
        for (size_t c = 0; c < ncell; c++) {
            if (!anom) {
                out[(size_t)L * c] = h[c];
                out[1 + (size_t)L * c] = h[ncell + c];
            }
            for (int b = 0; b < NBGC; b++)
                if (P.slot[b] >= 0) out[(size_t)P.slot[b] + (size_t)L * c] = h[(size_t)(anom ? b : P.ocn[b]) * ncell + c];
        }
    } else if (what == 1) {
        // after a coupling the particulates carry its salinity ratio, not yet
        // applied on the device (the next column step applies it when it reads them)
        std::vector<double> h((size_t)NSED * ncell);
        HIPCHK(e, hipMemcpy(h.data(), m->dev.part, h.size() * 8, hipMemcpyDeviceToHost));
        double ps = 1.0;
        if (!m->coupling_due) HIPCHK(e, hipMemcpy(&ps, m->dev.scal + SC_PSCALE, 8, hipMemcpyDeviceToHost));
        for (size_t c = 0; c < ncell; c++)
            for (int s = 0; s < NSED; s++)
                out[s + (size_t)NSED * c] = m->coupling_due ? h[(size_t)s * ncell + c] : ps * (h[(size_t)s * ncell + c] + 0.0);
    } else if (what == 2) {
        std::vector<double> h((size_t)NCARB * ncol);
        HIPCHK(e, hipMemcpy(h.data(), m->dev.carb, h.size() * 8, hipMemcpyDeviceToHost));
        for (size_t c = 0; c < ncol; c++)
            for (int s = 0; s < NCARB; s++) out[s + (size_t)NCARB * c] = h[(size_t)s * ncol + c];
    } else if (what == 5) {
        std::vector<double> h((size_t)NSED * ncol);
        HIPCHK(e, hipMemcpy(h.data(), m->dev.settle, h.size() * 8, hipMemcpyDeviceToHost));
        for (size_t c = 0; c < ncol; c++)
            for (int s = 0; s < NSED; s++) out[s + (size_t)NSED * c] = h[(size_t)s * ncol + c];
    } else if (what == 4) {
        int rc = sync_table(e);
        if (rc) return rc;
        std::vector<double> h((size_t)NCC_H * ncol, 0.0);
        double* d = nullptr;
        HIPCHK(e, hipMalloc(&d, h.size() * 8));
        hipMemset(d, 0, h.size() * 8);
        launch_carbconst_surface(e->d_table, id, g.nwet, d, e->s);
        hipError_t r = hipStreamSynchronize(e->s);
        if (r == hipSuccess) r = hipMemcpy(h.data(), d, h.size() * 8, hipMemcpyDeviceToHost);
        hipFree(d);
        if (r != hipSuccess) return fail(e, CGO_E_HIP, "bgc_download carbconst: %s", hipGetErrorString(r));
        for (size_t c = 0; c < ncol; c++)
            for (int s = 0; s < NCC_H; s++) out[s + (size_t)NCC_H * c] = h[(size_t)s * ncol + c];
    } else if (what == 6 || what == 7) {
        std::vector<double> h((size_t)2 * ncol);
        HIPCHK(e, hipMemcpy(h.data(), what == 6 ? m->dev.atm : m->dev.fatm, h.size() * 8, hipMemcpyDeviceToHost));
        for (size_t c = 0; c < ncol; c++)
            for (int q = 0; q < 2; q++) out[q + 2 * c] = h[(size_t)q * ncol + c];
    } else {
        return fail(e, CGO_E_ARG, "bgc_download: what %d", what);
    }
    return CGO_OK;
}

int cgo_set_windspeed(cgo_engine* e, int32_t id, const double* wspeed) {
    MemberH* m = member(e, id);
    if (!m || !wspeed || !m->bgc) return fail(e, CGO_E_ARG, "set_windspeed: bad argument or no BIOGEM state");
    const size_t ncol = m->dev.g.ncol;
    for (size_t c = 0; c < ncol; c++)
        if (!(wspeed[c] >= 0.0) || !std::isfinite(wspeed[c]))
            return fail(e, CGO_E_ARG, "set_windspeed: wind speed %g at column %zu", wspeed[c], c);
    if (int rc = settle(e, m)) return rc;
    HIPCHK(e, hipMemcpy(m->dev.wspeed, wspeed, ncol * 8, hipMemcpyHostToDevice));
    return CGO_OK;
}

// ---------------------------------------------------------------- stepping
struct Lists {
    // [p64 list][p32 list][bgc p64][bgc p32][all][velocity solve][surface fluxes]
    // [velocity solve through the band][velocity solve by the dense inverse, grouped by grid]
    std::vector<int32_t> v;
    int n[9] = {}, off[9] = {};
    std::vector<int> momGroups;             // per grid of the dense solve: offset in list 8, count, N
    int momN = 0, momCol = 0, momBand = 0;  // most psi points, columns and maxi of a member with a velocity solve
    int sfxCol = 0;                         // most columns of a member with surface fluxes
    int maxcol[2] = {0, 0}, maxL[2] = {0, 0}, maxwet[2] = {0, 0}, maxwet_bgc[2] = {0, 0}, maxi[2] = {0, 0};
    int maxk_bgc[2] = {0, 0}, wstream[2] = {0, 0}, maxk[2] = {0, 0};
    int ncol_atm = 0;  // widest grid of the members with air-sea exchange, 0 if none
    FluxShape fs[2] = {};
};

static int make_lists(cgo_engine* e, const int32_t* ids, int32_t n, Lists& ls) {
    std::vector<int32_t> g[9];
    std::vector<int> dgrid;                 // grids of the dense solve, in order of first appearance
    std::vector<std::vector<int32_t>> dmem;
    for (int i = 0; i < n; i++) {
        MemberH* m = member(e, ids[i]);
        if (!m) return fail(e, CGO_E_ARG, "step: bad member %d", ids[i]);
        const int p = m->prec == CGO_FP64 ? 0 : 1;
        g[p].push_back(ids[i]);
        if (m->bgc) g[2 + p].push_back(ids[i]);
        g[4].push_back(ids[i]);
        if (m->mbuf) {
            g[5].push_back(ids[i]);
            ls.momN = std::max(ls.momN, e->grids[m->grid]->mom->N);
            ls.momCol = std::max(ls.momCol, m->dev.g.ncol);
            ls.momBand = std::max(ls.momBand, m->dev.g.maxi);
            // FP32 members solve by the dense inverse where their grid keeps one
            if (m->prec != CGO_FP64 && e->grids[m->grid]->ainv_on) {
                size_t q = 0;
                while (q < dgrid.size() && dgrid[q] != m->grid) q++;
                if (q == dgrid.size()) {
                    dgrid.push_back(m->grid);
                    dmem.emplace_back();
                }
                dmem[q].push_back(ids[i]);
            } else {
                g[7].push_back(ids[i]);
            }
        }
        if (m->sbuf) {
            if (e->grids[m->grid]->sfx_atchem && !(m->bgc && m->dev.bp.airsea))
                return fail(e, CGO_E_STATE, "step: member %d's surface fluxes read CO2 from an atmosphere it does not have "
                                            "(cgo_bgc_init with air-sea exchange)", ids[i]);
            g[6].push_back(ids[i]);
            ls.sfxCol = std::max(ls.sfxCol, m->dev.g.ncol);
        }
        ls.maxcol[p] = std::max(ls.maxcol[p], m->dev.g.ncol);
        ls.maxL[p] = std::max(ls.maxL[p], m->L);
        ls.maxi[p] = std::max(ls.maxi[p], m->dev.g.maxi);
        ls.maxwet[p] = std::max(ls.maxwet[p], m->dev.g.nwet);
        ls.maxk[p] = std::max(ls.maxk[p], m->dev.g.maxk);
        for (int v = 0; v < 2; v++) {
            ls.fs[p].nblk[v] = std::max(ls.fs[p].nblk[v], m->dev.g.nfblk[v]);
            ls.fs[p].tw[v] = std::max(ls.fs[p].tw[v], e->grids[m->grid]->ftw[v]);
        }
        if (m->bgc) {
            ls.maxwet_bgc[p] = std::max(ls.maxwet_bgc[p], m->dev.g.nwet);
            ls.maxk_bgc[p] = std::max(ls.maxk_bgc[p], m->dev.g.maxk);
            ls.wstream[p] = std::max(ls.wstream[p], m->wstream);
            if (m->dev.bp.airsea) ls.ncol_atm = std::max(ls.ncol_atm, m->dev.g.ncol);
        }
    }
    for (size_t q = 0; q < dgrid.size(); q++) {
        ls.momGroups.push_back((int)g[8].size());
        ls.momGroups.push_back((int)dmem[q].size());
        ls.momGroups.push_back(e->grids[dgrid[q]]->mom->N);
        g[8].insert(g[8].end(), dmem[q].begin(), dmem[q].end());
    }
    for (int q = 0; q < 9; q++) {
        ls.off[q] = (int)ls.v.size();
        ls.n[q] = (int)g[q].size();
        ls.v.insert(ls.v.end(), g[q].begin(), g[q].end());
    }
    return CGO_OK;
}

// tstepo: flux + co per precision, then the parity flip (goldstein.f90:2280-2432)
// the flux launches of both precisions; part as launch_flux_t
static void enqueue_flux(cgo_engine* e, const Lists& ls, const int32_t* d_lists, int part) {
    for (int p = 0; p < 2; p++) {
        if (!ls.n[p]) continue;
        const int32_t* l = d_lists + ls.off[p];
        if (p == 0)
            launch_flux64(e->d_table, l, ls.n[p], ls.fs[p], ls.maxL[p], e->lb[0], e->nt[0], e->nbuf[0], e->isosplit[0],
                          part, e->lds_cap, e->s);
        else
            launch_flux32(e->d_table, l, ls.n[p], ls.fs[p], ls.maxL[p], e->lb[1], e->nt[1], e->nbuf[1], e->isosplit[1],
                          part, e->lds_cap, e->s);
    }
}

// co and the parity flip that end tstepo
static void enqueue_co_flip(cgo_engine* e, const Lists& ls, const int32_t* d_lists) {
    for (int p = 0; p < 2; p++)
        if (ls.n[p]) launch_co(e->d_table, d_lists + ls.off[p], ls.n[p], ls.maxwet[p], ls.maxk[p], p, e->s);
    launch_flip(e->d_table, d_lists + ls.off[4], ls.n[4], e->s);
}

// the velocity solve of the members that have one (step_goldstein 198-233),
// which precedes their tstepo
// part 1: the column kernel before the solve; 2: the rest; 3: both
static void enqueue_mom(cgo_engine* e, const Lists& ls, const int32_t* d_lists, hipStream_t s, int part = 3) {
    if (ls.n[5])
        launch_mom(e->d_table, d_lists + ls.off[5], ls.n[5], d_lists + ls.off[7], ls.n[7], d_lists + ls.off[8],
                   ls.momGroups.data(), (int)ls.momGroups.size() / 3, ls.momN, ls.momCol, ls.momBand, part, s);
}

// the surface fluxes of the members that have them (surflux, called by
// genie before the ocean step, genie.f90:271-282)
static void enqueue_sfx(cgo_engine* e, const Lists& ls, const int32_t* d_lists) {
    if (ls.n[6]) launch_sfx(e->d_table, d_lists + ls.off[6], ls.n[6], ls.sfxCol, 0, e->s);
}

static void enqueue_tstepo(cgo_engine* e, const Lists& ls, const int32_t* d_lists) {
    enqueue_mom(e, ls, d_lists, e->s);
    enqueue_flux(e, ls, d_lists, 3);
    enqueue_co_flip(e, ls, d_lists);
}

// step_biogem for the members with BIOGEM state (biogem.f90:528-1877), then
// for those with air-sea exchange the ocean -> atmosphere coupling and ATCHEM
// (genie.f90:410-450; they read only the flux and the atmosphere, which
// nothing else touches, so they may run before the tracer coupling)
static void enqueue_bio(cgo_engine* e, const Lists& ls, const int32_t* d_lists, hipStream_t s) {
    for (int p = 0; p < 2; p++)
        if (ls.n[2 + p])
            launch_bio(e->d_table, d_lists + ls.off[2 + p], ls.n[2 + p], ls.maxwet_bgc[p], ls.wstream[p], p, s);
    if (ls.ncol_atm)
        for (int p = 0; p < 2; p++)
            if (ls.n[2 + p]) launch_atchem(e->d_table, d_lists + ls.off[2 + p], ls.n[2 + p], ls.ncol_atm, s);
}

// biogem_tracercoupling (biogem.f90:1885-2077); phase as launch_coupling
static void enqueue_coupling(cgo_engine* e, const Lists& ls, const int32_t* d_lists, int phase = 0) {
    for (int p = 0; p < 2; p++)
        if (ls.n[2 + p])
            launch_coupling(e->d_table, d_lists + ls.off[2 + p], ls.n[2 + p], ls.maxwet_bgc[p], ls.maxk_bgc[p],
                            ls.maxL[p], p, e->lds_cap, phase, e->s);
}

// Whole ocean steps in the genie.f90 order.  step_biogem reads only BIOGEM's
// own arrays, which the previous coupling wrote, and writes none that the
// transport step or the coupling's sums read: it runs on a second stream,
// concurrently with tstepo and the global sums, and joins before the rescale.
static void enqueue_steps(cgo_engine* e, const Lists& ls, const int32_t* d_lists, int nsteps) {
    const bool bgc = ls.n[2] || ls.n[3];
    for (int it = 0; it < nsteps; it++) {
        // the BIOGEM step depends only on the last coupling: it forks from
        // there and runs beside the isoneutral pre-pass and the flux's tracer
        // blocks.  Its launch is enqueued after the pre-pass's, which a
        // captured graph keeps: the pre-pass (short, latency-bound) takes the
        // CUs first, and BIOGEM's blocks fill the slots both flux kernels
        // leave (enqueued first, they ran beside the pre-pass alone)
        // the surface fluxes read the atmosphere that the last step's ATCHEM
        // left: they come before the fork, so this step's ATCHEM waits for them.
        // The velocity solve neither reads nor writes BIOGEM's arrays; the
        // fork follows its first column kernel, so the BIOGEM step runs beside
        // the solve blocks (one per member) and the small column kernels
        // after them rather than beside that column kernel
        enqueue_sfx(e, ls, d_lists);
        enqueue_mom(e, ls, d_lists, e->s, 1);
        if (bgc) hipEventRecord(e->ev_fork, e->s);
        enqueue_mom(e, ls, d_lists, e->s, 2);
        enqueue_flux(e, ls, d_lists, 1);
        if (bgc) {
            hipStreamWaitEvent(e->s2, e->ev_fork, 0);
            enqueue_bio(e, ls, d_lists, e->s2);
            hipEventRecord(e->ev_join, e->s2);
        }
        enqueue_flux(e, ls, d_lists, 2);
        enqueue_co_flip(e, ls, d_lists);
        if (bgc) {
            enqueue_coupling(e, ls, d_lists, 3);  // the sums are fresh: tstepo formed them
            hipStreamWaitEvent(e->s, e->ev_join, 0);
            enqueue_coupling(e, ls, d_lists, 2);
        }
    }
}

// a member between cgo_bio_step and cgo_tracercoupling accepts only the coupling
static int check_not_due(cgo_engine* e, const int32_t* ids, int32_t n, const char* who) {
    for (int i = 0; i < n; i++) {
        MemberH* m = member(e, ids[i]);
        if (m && m->coupling_due)
            return fail(e, CGO_E_STATE, "%s: member %d ran cgo_bio_step without cgo_tracercoupling", who, ids[i]);
    }
    return CGO_OK;
}

static int upload_lists(cgo_engine* e, const Lists& ls) {
    if (ls.v == e->last_lists && e->d_list) return CGO_OK;
    if (ls.v.size() > e->d_list_cap) {
        if (e->d_list) hipFree(e->d_list);
        e->d_list_cap = std::max<size_t>(ls.v.size(), 1024);
        HIPCHK(e, hipMalloc(&e->d_list, e->d_list_cap * 4));
    }
    HIPCHK(e, hipStreamSynchronize(e->s));
    HIPCHK(e, hipMemcpy(e->d_list, ls.v.data(), ls.v.size() * 4, hipMemcpyHostToDevice));
    e->last_lists = ls.v;
    return CGO_OK;
}

static int prepare(cgo_engine* e, const int32_t* ids, int32_t n, Lists& ls, bool steps = true) {
    hipSetDevice(e->device);
    int rc = make_lists(e, ids, n, ls);
    if (rc) return rc;
    if ((rc = sync_table(e))) return rc;
    if ((rc = upload_lists(e, ls))) return rc;
    for (int i = 0; i < n; i++) {
        member(e, ids[i])->queued = true;
        if (steps) member(e, ids[i])->stepped = true;
    }
    return CGO_OK;
}

int cgo_step(cgo_engine* e, const int32_t* ids, int32_t n, int32_t nsteps) {
    if (!e || !ids || n <= 0 || nsteps < 0) return fail(e, CGO_E_ARG, "step: bad argument");
    Lists ls;
    int rc = check_not_due(e, ids, n, "step");
    if (rc || (rc = prepare(e, ids, n, ls))) return rc;
    enqueue_steps(e, ls, e->d_list, nsteps);
    HIPCHK(e, hipGetLastError());
    return CGO_OK;
}

int cgo_surflux(cgo_engine* e, const int32_t* ids, int32_t n) {
    if (!e || !ids || n <= 0) return fail(e, CGO_E_ARG, "surflux: bad argument");
    for (int i = 0; i < n; i++) {
        MemberH* m = member(e, ids[i]);
        if (!m) return fail(e, CGO_E_ARG, "surflux: bad member %d", ids[i]);
        if (!m->sbuf) return fail(e, CGO_E_STATE, "surflux: member %d has no surface fluxes (cgo_set_surface)", ids[i]);
    }
    Lists ls;
    if (int rc = prepare(e, ids, n, ls, false)) return rc;
    enqueue_sfx(e, ls, e->d_list);
    HIPCHK(e, hipGetLastError());
    return CGO_OK;
}

int cgo_surface_slot(cgo_engine* e, int32_t id, const double* fx0sic, const double* fwsic) {
    MemberH* m = member(e, id);
    if (!m) return fail(e, CGO_E_ARG, "surface_slot: bad member %d", id);
    if (!m->sbuf) return fail(e, CGO_E_STATE, "surface_slot: member %d has no surface fluxes (cgo_set_surface)", id);
    if (int rc = settle(e, m)) return rc;
    const size_t ncol = (size_t)m->dev.g.ncol;
    std::vector<double> h(2 * ncol, 0.0);
    if (fx0sic) std::copy(fx0sic, fx0sic + ncol, h.begin());
    if (fwsic) std::copy(fwsic, fwsic + ncol, h.begin() + ncol);
    HIPCHK(e, hipMemcpy(static_cast<double*>(m->sbuf) + SW_FX0SIC * ncol, h.data(), h.size() * 8, hipMemcpyHostToDevice));
    Lists ls;
    if (int rc = prepare(e, &id, 1, ls, false)) return rc;
    launch_sfx(e->d_table, e->d_list + ls.off[6], ls.n[6], ls.sfxCol, 1, e->s);
    HIPCHK(e, hipGetLastError());
    return CGO_OK;
}

int cgo_tstepo(cgo_engine* e, const int32_t* ids, int32_t n, int32_t nsteps) {
    if (!e || !ids || n <= 0 || nsteps < 0) return fail(e, CGO_E_ARG, "tstepo: bad argument");
    Lists ls;
    int rc = check_not_due(e, ids, n, "tstepo");
    if (rc || (rc = prepare(e, ids, n, ls))) return rc;
    for (int it = 0; it < nsteps; it++) enqueue_tstepo(e, ls, e->d_list);
    HIPCHK(e, hipGetLastError());
    return CGO_OK;
}

int cgo_bio_step(cgo_engine* e, const int32_t* ids, int32_t n) {
    if (!e || !ids || n <= 0) return fail(e, CGO_E_ARG, "bio_step: bad argument");
    for (int i = 0; i < n; i++) {
        MemberH* m = member(e, ids[i]);
        if (!m) return fail(e, CGO_E_ARG, "bio_step: bad member %d", ids[i]);
        if (!m->bgc) return fail(e, CGO_E_STATE, "bio_step: member %d has no BIOGEM state (cgo_bgc_init)", ids[i]);
    }
    Lists ls;
    int rc = check_not_due(e, ids, n, "bio_step");
    if (rc || (rc = prepare(e, ids, n, ls))) return rc;
    enqueue_bio(e, ls, e->d_list, e->s);
    HIPCHK(e, hipGetLastError());
    for (int i = 0; i < n; i++) member(e, ids[i])->coupling_due = true;
    return CGO_OK;
}

int cgo_tracercoupling(cgo_engine* e, const int32_t* ids, int32_t n) {
    if (!e || !ids || n <= 0) return fail(e, CGO_E_ARG, "tracercoupling: bad argument");
    for (int i = 0; i < n; i++) {
        MemberH* m = member(e, ids[i]);
        if (!m) return fail(e, CGO_E_ARG, "tracercoupling: bad member %d", ids[i]);
        if (!m->coupling_due)
            return fail(e, CGO_E_STATE, "tracercoupling: member %d needs cgo_bio_step first", ids[i]);
    }
    Lists ls;
    int rc = prepare(e, ids, n, ls);
    if (rc) return rc;
    enqueue_coupling(e, ls, e->d_list);
    HIPCHK(e, hipGetLastError());
    for (int i = 0; i < n; i++) member(e, ids[i])->coupling_due = false;
    return CGO_OK;
}

int cgo_graph_capture(cgo_engine* e, const int32_t* ids, int32_t n, int32_t nsteps, int32_t* graph_id) {
    if (!e || !ids || n <= 0 || nsteps <= 0 || !graph_id) return fail(e, CGO_E_ARG, "graph_capture: bad argument");
    hipSetDevice(e->device);
    Lists ls;
    int rc = check_not_due(e, ids, n, "graph_capture");
    if (rc) return rc;
    rc = make_lists(e, ids, n, ls);
    if (rc) return rc;
    if ((rc = sync_table(e))) return rc;
    Graph gr;
    gr.ids.assign(ids, ids + n);
    for (int i = 0; i < n; i++) gr.gens.push_back(ids[i] < (int32_t)e->gen.size() ? e->gen[ids[i]] : 0);
    HIPCHK(e, hipMalloc(&gr.d_lists, ls.v.size() * 4));
    HIPCHK(e, hipMemcpy(gr.d_lists, ls.v.data(), ls.v.size() * 4, hipMemcpyHostToDevice));
    HIPCHK(e, hipStreamSynchronize(e->s));
    HIPCHK(e, hipStreamBeginCapture(e->s, hipStreamCaptureModeThreadLocal));
    enqueue_steps(e, ls, gr.d_lists, nsteps);
    hipError_t r = hipStreamEndCapture(e->s, &gr.graph);
    if (r != hipSuccess) {
        hipFree(gr.d_lists);
        return fail(e, CGO_E_HIP, "hipStreamEndCapture: %s", hipGetErrorString(r));
    }
    r = hipGraphInstantiate(&gr.exec, gr.graph, nullptr, nullptr, 0);
    if (r != hipSuccess) {
        hipGraphDestroy(gr.graph);
        hipFree(gr.d_lists);
        return fail(e, CGO_E_HIP, "hipGraphInstantiate: %s", hipGetErrorString(r));
    }
    int slot = -1;
    for (size_t i = 0; i < e->graphs.size(); i++)
        if (!e->graphs[i].exec) slot = (int)i;
    if (slot < 0) {
        e->graphs.push_back(gr);
        slot = (int)e->graphs.size() - 1;
    } else {
        e->graphs[slot] = gr;
    }
    *graph_id = slot;
    return CGO_OK;
}

int cgo_graph_launch(cgo_engine* e, int32_t gid, int32_t nreps) {
    if (!e || gid < 0 || gid >= (int)e->graphs.size() || !e->graphs[gid].exec) return fail(e, CGO_E_ARG, "graph_launch: bad graph");
    if (nreps < 0) return fail(e, CGO_E_ARG, "graph_launch: nreps %d", nreps);
    const Graph& gr = e->graphs[gid];
    // every captured member must still be the one captured (same slot
    // generation: not destroyed, re-created or re-laid-out since) and must
    // not be between cgo_bio_step and cgo_tracercoupling
    for (size_t i = 0; i < gr.ids.size(); i++) {
        const int32_t id = gr.ids[i];
        if (!member(e, id) || id >= (int32_t)e->gen.size() || e->gen[id] != gr.gens[i])
            return fail(e, CGO_E_STATE, "graph_launch: member %d was destroyed, re-created or re-initialised since the capture; capture again", id);
    }
    int rc = check_not_due(e, gr.ids.data(), (int32_t)gr.ids.size(), "graph_launch");
    if (rc) return rc;
    hipSetDevice(e->device);
    // the graph reads the member table at run time: bring it up to date
    // (e.g. a surface-flux buffer attached since the capture)
    if ((rc = sync_table(e))) return rc;
    for (int32_t id : gr.ids) {
        member(e, id)->queued = true;
        member(e, id)->stepped = true;
    }
    for (int r = 0; r < nreps; r++) HIPCHK(e, hipGraphLaunch(gr.exec, e->s));
    return CGO_OK;
}

int cgo_graph_destroy(cgo_engine* e, int32_t gid) {
    if (!e || gid < 0 || gid >= (int)e->graphs.size() || !e->graphs[gid].exec) return fail(e, CGO_E_ARG, "graph_destroy: bad graph");
    HIPCHK(e, hipStreamSynchronize(e->s));
    Graph& g = e->graphs[gid];
    hipGraphExecDestroy(g.exec);
    hipGraphDestroy(g.graph);
    hipFree(g.d_lists);
    g = Graph();
    return CGO_OK;
}

int cgo_set_tile(cgo_engine* e, int32_t precision, int32_t tb) {
    if (!e || (precision != CGO_FP64 && precision != CGO_FP32C) || (tb != 2 && tb != 4 && tb != 6 && tb != 8))
        return fail(e, CGO_E_ARG, "set_tile: precision %d tracer block %d", precision, tb);
    e->lb[precision == CGO_FP64 ? 0 : 1] = tb;
    return CGO_OK;
}

int cgo_set_flux_cols(cgo_engine* e, int32_t precision, int32_t cols) {
    if (!e || (precision != CGO_FP64 && precision != CGO_FP32C) || (cols != 128 && cols != 256))
        return fail(e, CGO_E_ARG, "set_flux_cols: precision %d columns %d", precision, cols);
    e->nt[precision == CGO_FP64 ? 0 : 1] = cols;
    return CGO_OK;
}

int cgo_set_flux_ring(cgo_engine* e, int32_t precision, int32_t planes) {
    if (!e || (precision != CGO_FP64 && precision != CGO_FP32C) || (planes != 2 && planes != 3))
        return fail(e, CGO_E_ARG, "set_flux_ring: precision %d planes %d", precision, planes);
    e->nbuf[precision == CGO_FP64 ? 0 : 1] = planes;
    return CGO_OK;
}

int cgo_set_flux_iso(cgo_engine* e, int32_t precision, int32_t prepass) {
    if (!e || (precision != CGO_FP64 && precision != CGO_FP32C) || (prepass != 0 && prepass != 1))
        return fail(e, CGO_E_ARG, "set_flux_iso: precision %d prepass %d", precision, prepass);
    e->isosplit[precision == CGO_FP64 ? 0 : 1] = prepass;
    return CGO_OK;
}

int cgo_time_kernel(cgo_engine* e, int32_t which, const int32_t* ids, int32_t n, int32_t reps, double* avg_ms) {
    if (!e || !ids || n <= 0 || reps <= 0 || !avg_ms || which < 0 || which > 6) return fail(e, CGO_E_ARG, "time_kernel: bad argument");
    Lists ls;
    int rc = make_lists(e, ids, n, ls);
    if (rc) return rc;
    if ((rc = sync_table(e))) return rc;
    if ((rc = upload_lists(e, ls))) return rc;
    HIPCHK(e, hipStreamSynchronize(e->s));
    HIPCHK(e, hipEventRecord(e->ev0, e->s));
    if (which == 3) {
        // the flux as it runs inside a whole step: the isoneutral pre-pass and
        // the tracer blocks with the BIOGEM column step on the second stream,
        // enqueued as a step enqueues them
        const bool bgc = ls.n[2] || ls.n[3];
        for (int r = 0; r < reps; r++) {
            if (bgc) hipEventRecord(e->ev_fork, e->s);
            enqueue_flux(e, ls, e->d_list, 1);
            if (bgc) {
                hipStreamWaitEvent(e->s2, e->ev_fork, 0);
                enqueue_bio(e, ls, e->d_list, e->s2);
                hipEventRecord(e->ev_join, e->s2);
            }
            enqueue_flux(e, ls, e->d_list, 2);
            if (bgc) hipStreamWaitEvent(e->s, e->ev_join, 0);
        }
    }
    if (which == 5)
        for (int r = 0; r < reps; r++) enqueue_mom(e, ls, e->d_list, e->s);
    if (which == 6)
        for (int r = 0; r < reps; r++) enqueue_sfx(e, ls, e->d_list);
    for (int r = 0; r < (which == 3 || which >= 5 ? 0 : reps); r++)
        for (int p = 0; p < 2; p++) {
            if (!ls.n[p]) continue;
            const int32_t* l = e->d_list + ls.off[p];
            if (which == 0) {
                if (p == 0)
                    launch_flux64(e->d_table, l, ls.n[p], ls.fs[p], ls.maxL[p], e->lb[0], e->nt[0], e->nbuf[0], e->isosplit[0], 3, e->lds_cap,
                          e->s);
                else
                    launch_flux32(e->d_table, l, ls.n[p], ls.fs[p], ls.maxL[p], e->lb[1], e->nt[1], e->nbuf[1], e->isosplit[1], 3, e->lds_cap,
                          e->s);
            } else if (which == 1) {
                launch_co(e->d_table, l, ls.n[p], ls.maxwet[p], ls.maxk[p], p, e->s);
            } else if (which == 4) {
                if (ls.n[2 + p])
                    launch_bio(e->d_table, e->d_list + ls.off[2 + p], ls.n[2 + p], ls.maxwet_bgc[p], ls.wstream[p], p, e->s);
            } else if (ls.n[2 + p]) {
                launch_bio(e->d_table, e->d_list + ls.off[2 + p], ls.n[2 + p], ls.maxwet_bgc[p], ls.wstream[p], p, e->s);
                launch_coupling(e->d_table, e->d_list + ls.off[2 + p], ls.n[2 + p], ls.maxwet_bgc[p], ls.maxk_bgc[p],
                                ls.maxL[p], p, e->lds_cap, 0, e->s);
            }
        }
    HIPCHK(e, hipEventRecord(e->ev1, e->s));
    HIPCHK(e, hipEventSynchronize(e->ev1));
    float ms = 0;
    HIPCHK(e, hipEventElapsedTime(&ms, e->ev0, e->ev1));
    *avg_ms = (double)ms / reps;
    return CGO_OK;
}

}  // extern "C"
