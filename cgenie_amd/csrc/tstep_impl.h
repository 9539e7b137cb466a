// tstep_impl.h -- GOLDSTEIN tracer step kernels (tstepo_flux + eos, co).
//
// Included by k_tstep_f64.hip (built with -ffp-contract=off: every operation
// rounds as the reference writes it, so the FP64 path reproduces the CPU
// restatement bit for bit) and by k_tstep_f32.hip (the FP32 form).
//
// k_flux: one thread per (member, water column, tracer block).  The thread
// marches k upward like goldstein.f90:2466-2641 and carries the bottom face
// flux fb from the level below (2635), so every vertical face flux is
// computed once.  Horizontal face fluxes are evaluated by both neighbours
// with the same expression and operands (the east flux of cell i is the
// west flux of cell i+1, 2633), so both sides see the same bits and the
// scheme stays conservative.  Level k+1 values are loaded once and carried
// down to level k of the next iteration.
#pragma once
#include "cgocn_dev.h"

namespace cgo {


template <typename R> struct Met;
template <> struct Met<double> {
    static __device__ __forceinline__ const double* j(const GridDev& g) { return g.mj; }
    static __device__ __forceinline__ const double* k(const GridDev& g) { return g.mk; }
};
template <> struct Met<float> {
    static __device__ __forceinline__ const float* j(const GridDev& g) { return g.mjf; }
    static __device__ __forceinline__ const float* k(const GridDev& g) { return g.mkf; }
};

// weight of one unit of the int16 carry of a float value x: 2^-16 ulp(x)
__device__ __forceinline__ double carry_unit(float x) { return ldexp(1.0, carry_exp(x) - 23 - 16); }

// state access as double (FP32 form: hi + carry)
// a / b: the IEEE quotient in FP64; in the FP32 form the hardware reciprocal
// (1 ulp), for the per-cell isoneutral coefficients
__device__ __forceinline__ double fdiv(double a, double b) { return a / b; }
__device__ __forceinline__ float fdiv(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }

template <typename R> struct StateIO;
// (all state pointers are global memory: the casts keep the accesses off
// the flat path, whose loads also count against the LDS wait counter)
// ldr/val split a load from its conversion so that a thread can issue the
// loads of many cells before it waits for any of them.
template <> struct StateIO<double> {
    struct Raw {
        double v;
    };
    static __device__ __forceinline__ Raw ldr(const void* s, const int16_t*, uint32_t i) {
        return Raw{at((const AS_GLOBAL double*)s, i)};
    }
    static __device__ __forceinline__ double val(Raw r, bool) { return r.v; }
    static __device__ __forceinline__ double ld(const void* s, const int16_t*, uint32_t i) {
        return at((const AS_GLOBAL double*)s, i);
    }
    static __device__ __forceinline__ void st(void* s, int16_t*, uint32_t i, double x) {
        at((AS_GLOBAL double*)s, i) = x;
    }
    // the value a later ld returns after st(x)
    static __device__ __forceinline__ double rt(double x, bool) { return x; }
};
template <> struct StateIO<float> {
    struct Raw {
        float h;
        int16_t c;
    };
    static __device__ __forceinline__ Raw ldr(const void* s, const int16_t* lo, uint32_t i) {
        Raw r;
        r.h = at((const AS_GLOBAL float*)s, i);
        r.c = lo ? at((const AS_GLOBAL int16_t*)lo, i) : (int16_t)0;
        return r;
    }
    static __device__ __forceinline__ double val(Raw r, bool carry) {
        double x = (double)r.h;
        if (carry) x += (double)r.c * carry_unit(r.h);
        return x;
    }
    static __device__ __forceinline__ double ld(const void* s, const int16_t* lo, uint32_t i) {
        float h = at((const AS_GLOBAL float*)s, i);
        double x = (double)h;
        if (lo) x += (double)at((const AS_GLOBAL int16_t*)lo, i) * carry_unit(h);
        return x;
    }
    static __device__ __forceinline__ void st(void* s, int16_t* lo, uint32_t i, double x) {
        float h = (float)x;
        at((AS_GLOBAL float*)s, i) = h;
        if (lo) {
            // x - h in units of carry_unit(h): scaling by the exact power of two
            // 2^(39 - e) rounds exactly as dividing by 2^(e - 39) would; the
            // rounded value is within +-2^15, so the clamp is taken on the integer
            const double r = rint(ldexp(x - (double)h, 23 + 16 - carry_exp(h)));
            int ri = (int)r;
            ri = ri < -32767 ? -32767 : (ri > 32767 ? 32767 : ri);
            at((AS_GLOBAL int16_t*)lo, i) = (int16_t)ri;
        }
    }
    // the value a later ld returns after st(x): the float, plus the carry
    // split and clamped as st stores it
    static __device__ __forceinline__ double rt(double x, bool carry) {
        const float h = (float)x;
        double v = (double)h;
        if (carry) {
            const double r = rint(ldexp(x - (double)h, 23 + 16 - carry_exp(h)));
            int ri = (int)r;
            ri = ri < -32767 ? -32767 : (ri > 32767 ? 32767 : ri);
            v += (double)(int16_t)ri * carry_unit(h);
        }
        return v;
    }
};

__device__ __forceinline__ void atomic_max_pos(double* p, double v) {
    // non-negative doubles order like their bit patterns
    atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

// Work of k_flux: a block takes up to FLUX_THREADS consecutive wet columns
// (ascending column order; land columns get no thread) and the tracer block
// [l0, l0 + LB).  LDS layout: the k1 tile, a ring of NBUF level planes of LB
// tracer slots plus T and S, each covering the block's flattened column range
// [t0, t1) -- every east/west (periodic wrap included, offsets below maxi)
// and north/south neighbour of its columns -- and the carried bottom-face
// flux.  With NBUF = 3 the planes k, k+1 and the prefetched k+2 are
// resident; with NBUF = 2 the prefetched plane replaces plane k after a
// second barrier per level (less LDS, more workgroups per CU).  The host cuts
// blocks so that t1 - t0 <= flux_tw(FLUX_THREADS).
// The FP32 form also stages the carry of every value, as a float residual,
// so differences of neighbours are taken at the stored (~40-bit) precision.
// ns level-plane slots (lb tracers, plus T and S when the block computes the
// isoneutral coefficients itself), lb carried bottom fluxes.  Every slot
// holds flux_tw(nt) positions, a compile-time stride, so each slot's offset
// is an immediate of the LDS instruction.
__host__ __device__ constexpr size_t flux_lds_bytes(int ns, int lb, int es, int nt, int nbuf) {
    return (size_t)flux_tw(nt) * 4 + (size_t)nbuf * ns * flux_tw(nt) * es + (size_t)lb * nt * es +
           (es == 4 ? (size_t)nbuf * ns * flux_tw(nt) * 4 : 0);
}

// the carry c of a float h as a float: c * 2^-16 ulp(h) (exact unless |h| < 2^-87)
// (c 2^-39 is exact; times 2^e, the exponent bits of h, it rounds once, as
// ldexpf(c, e - 39) does; zero and subnormal h give 0, as ldexpf(c, -165) does)
__device__ __forceinline__ float carry_residual(float h, int16_t c) {
    return ((float)c * 0x1p-39f) * __uint_as_float(__float_as_uint(h) & 0x7f800000u);
}

// IM selects where the isoneutral coefficients come from: 0 computed in every
// tracer block from T and S staged beside its tracers; 1 the pre-pass (one
// block per column block, T and S only, LB = 2) that computes them once per
// cell into the member's isoc planes; 2 the tracer blocks after that pre-pass,
// which load them and stage only their own tracers.
template <typename R, int LB, int FLUX_THREADS, int FLUX_NBUF, int IM>
__global__ __launch_bounds__(FLUX_THREADS) void k_flux(const MemberDev* __restrict__ mems,
                                                       const int32_t* __restrict__ list) {
    constexpr int NS = IM == 0 ? LB + 2 : LB;  // slots per staged level plane
    constexpr int TWS = flux_tw(FLUX_THREADS);  // positions per slot (>= every block's t1 - t0)
    extern __shared__ __align__(16) unsigned char smem[];
    const BlockXYZ B = xcd_blk();
    const MemberDev& M = mems[list[B.z]];
    const GridDev& g = M.g;
    const int maxi = g.maxi, maxj = g.maxj, maxk = g.maxk, L = M.L;
    constexpr int V = FLUX_THREADS == 256 ? 1 : 0;
    const int l0 = B.y * LB;
    if (B.x >= g.nfblk[V] || l0 >= L) return;  // uniform for the block
    // every wet column's sums are formed below
    if (IM != 1 && M.bgc && B.x == 0 && B.y == 0 && threadIdx.x == 0) *gp(M.sfresh) = 1;
    const int nl = (L - l0) < LB ? (L - l0) : LB;
    const AS_GLOBAL FluxBlk* blk = gp(g.fblk[V]) + B.x;
    const int w0 = blk->w0, t0 = blk->t0, TW = blk->t1 - t0, nw = blk->w1 - w0;
    int* sk1 = reinterpret_cast<int*>(smem);
    R* sp = reinterpret_cast<R*>(smem + (size_t)TWS * 4);
    R* sfb = sp + FLUX_NBUF * NS * TWS;
    float* sr = reinterpret_cast<float*>(sfb + LB * FLUX_THREADS);  // FP32 form: residual ring
    // slot q < LB: tracer l0+q; T and S sit in slots 0/1 when l0 == 0, else
    // (coefficients computed here) LB/LB+1
    const int qT = (l0 == 0 || IM != 0) ? 0 : LB, qS = (l0 == 0 || IM != 0) ? 1 : LB + 1;
    const int nslot = (l0 == 0 || IM != 0) ? nl : LB + 2;  // slots staged per level
    __shared__ int s_kmin;
    const int tid = threadIdx.x;
    if (tid == 0) s_kmin = maxk + 1;
    for (int pp = tid; pp < TW; pp += FLUX_THREADS) {
        const int c = t0 + pp;
        sk1[pp] = k1at(g, c % maxi + 1, c / maxi + 1);
    }
    __syncthreads();
    const bool active = tid < nw;
    const int col = active ? gp(g.wetsorted)[w0 + tid] : t0;
    const int i0 = col % maxi, j0 = col / maxi;
    const int I = i0 + 1, J = j0 + 1;
    const int k1c = active ? k1at(g, I, J) : maxk + 1;
    const int k1e = k1at(g, I + 1, J), k1w = k1at(g, I - 1, J);
    const int k1n = k1at(g, I, J + 1), k1s = k1at(g, I, J - 1);
    if (k1c <= maxk) atomicMin(&s_kmin, k1c);
    __syncthreads();
    const int kmin = s_kmin;
    if (kmin > maxk) return;  // every column of the block is land

    constexpr bool ANOM = sizeof(R) == 4;  // FP32 form: fluxes of q - q_cell plus q_cell * div(u)
    const int par = *gp(M.parity);
    const AS_GLOBAL R* __restrict__ src = gp(static_cast<const R*>(M.st[par]));
    AS_GLOBAL R* __restrict__ dst = gp(static_cast<R*>(M.st[par ^ 1]));
    const AS_GLOBAL int16_t* __restrict__ lo_src = M.lo[par] ? gp(M.lo[par]) : nullptr;
    int16_t* __restrict__ lo_dst = M.lo[par ^ 1];
    const AS_GLOBAL R* __restrict__ u = gp(static_cast<const R*>(M.u));
    AS_GLOBAL double* __restrict__ rhop = gp(M.rho);
    const int plane = g.ncol, tstride = maxk * g.ncol, ncell = g.ncell;
    // neighbour positions in the tile
    const int p = col - t0;
    const int pe = p + (i0 + 1 == maxi ? 1 - maxi : 1);
    const int pw = p + (i0 == 0 ? maxi - 1 : -1);
    const int pn = p + (j0 + 1 < maxj ? maxi : 0);
    const int ps = p - (j0 > 0 ? maxi : 0);
    const int cw = j0 * maxi + (i0 == 0 ? maxi - 1 : i0 - 1);
    const int cs = (j0 > 0 ? j0 - 1 : 0) * maxi + i0;
    const AS_GLOBAL R* mj = gp(Met<R>::j(g));
    const AS_GLOBAL R* mk = gp(Met<R>::k(g));
    const int JS = maxj + 2, KS = maxk + 2;
    const R rc = mj[MJ_RC * JS + J], rc2 = mj[MJ_RC2 * JS + J];
    const R cvn = mj[MJ_CV * JS + J], cv2n = mj[MJ_CV2 * JS + J];
    const R cvs = mj[MJ_CV * JS + J - 1], cv2s = mj[MJ_CV2 * JS + J - 1];
    const R dsvn = mj[MJ_DSVM * JS + J], dsvs = mj[MJ_DSVM * JS + J - 1];
    const R rds = mj[MJ_RDS * JS + J];
    const R rdsvn = mj[MJ_RDSV * JS + J], rdsvs = mj[MJ_RDSV * JS + J - 1];
    const R dphi = (R)g.dphi, rdphi = (R)g.rdphi;
    const R diff1 = sizeof(R) == 8 ? (R)M.diff1 : (R)M.diff1f;
    const R diffv = sizeof(R) == 8 ? (R)M.diff2 : (R)M.diff2f;
    const R e0 = (R)g.ec[0], e1 = (R)g.ec[1], e2 = (R)g.ec[2], e3 = (R)g.ec[3];
    const R scc = M.diso ? e1 : (R)0;

    auto SP = [&](int buf, int q, int pp) -> R& { return sp[(buf * NS + q) * TWS + pp]; };
    auto SR = [&](int buf, int q, int pp) -> float& { return sr[(buf * NS + q) * TWS + pp]; };
    // difference of two staged values; the FP32 form adds the residuals
    auto D = [&](int buf, int q, int pa, int pb) -> R {
        if (ANOM) return (SP(buf, q, pa) - SP(buf, q, pb)) + (R)(SR(buf, q, pa) - SR(buf, q, pb));
        return SP(buf, q, pa) - SP(buf, q, pb);
    };
    // difference across levels (same column)
    auto DK = [&](int bu, int bc, int q, int pp) -> R {
        if (ANOM) return (SP(bu, q, pp) - SP(bc, q, pp)) + (R)(SR(bu, q, pp) - SR(bc, q, pp));
        return SP(bu, q, pp) - SP(bc, q, pp);
    };
    auto FB = [&](int q) -> R& { return sfb[q * FLUX_THREADS + tid]; };
    auto slot_tracer = [&](int q) { return q < LB ? l0 + q : q - LB; };
    // prefetch of one plane into registers, then into the LDS ring
    constexpr int FLUX_PF = flux_pf(FLUX_THREADS);
    R pf[NS][FLUX_PF];
    int16_t pc[ANOM ? NS : 1][FLUX_PF];
    auto fetch = [&](int kz) {
        const int o = kz * plane + t0;
#pragma unroll
        for (int r = 0; r < FLUX_PF; r++) {
            const int pp = tid + r * FLUX_THREADS;
            const bool ok = pp < TW && kz + 1 >= sk1[pp < TW ? pp : 0];
#pragma unroll
            for (int q = 0; q < NS; q++)
                if (q < nslot || (IM == 0 && q >= LB)) {
                    const bool ld = ok && (q < nslot);
                    const uint32_t i = (uint32_t)(slot_tracer(q) * tstride + o + pp);
                    pf[q][r] = ld ? at(src, i) : (R)0;
                    if (ANOM) pc[ANOM ? q : 0][r] = (ld && lo_src) ? at(lo_src, i) : (int16_t)0;
                }
        }
    };
    auto commit = [&](int buf) {
#pragma unroll
        for (int r = 0; r < FLUX_PF; r++) {
            const int pp = tid + r * FLUX_THREADS;
            if (pp < TW) {
#pragma unroll
                for (int q = 0; q < NS; q++)
                    if (q < nslot) {
                        SP(buf, q, pp) = pf[q][r];
                        if (ANOM) SR(buf, q, pp) = carry_residual((float)pf[q][r], pc[ANOM ? q : 0][r]);
                    }
            }
        }
    };

    for (int q = 0; q < LB; q++) FB(q) = (R)0;
    int32_t limps = 0;
    double dmaxv = 0.0;
    // The isoneutral slope terms are formed in FP64 in both forms.  The
    // density gradients they divide by are differences of nearly equal
    // terms, so the FP32 form takes its staged values with their residuals
    // and the FP64 metrics, and rounds the finished coefficients to float
    // once; in float they lose up to 1e-4 of their value where the
    // stratification is weak.  The FP64 form evaluates the restatement's
    // expressions on the same values, bit for bit.
    const double rcD = gp(g.mj)[MJ_RC * JS + J], cvnD = gp(g.mj)[MJ_CV * JS + J];
    const double cvsD = gp(g.mj)[MJ_CV * JS + J - 1];
    const double rdsvnD = gp(g.mj)[MJ_RDSV * JS + J], rdsvsD = gp(g.mj)[MJ_RDSV * JS + J - 1];
    const double rdphiD = g.rdphi, diff1D = M.diff1;
    const double e0D = g.ec[0], e1D = g.ec[1], e2D = g.ec[2], e3D = g.ec[3];
    const double sccD = M.diso ? e1D : 0.0;
    auto SPD = [&](int buf, int q, int pp) -> double {
        if (ANOM) return (double)SP(buf, q, pp) + (double)SR(buf, q, pp);
        return (double)SP(buf, q, pp);
    };
    auto DD = [&](int buf, int q, int pa, int pb) -> double {
        if (ANOM)
            return ((double)SP(buf, q, pa) - (double)SP(buf, q, pb)) + ((double)SR(buf, q, pa) - (double)SR(buf, q, pb));
        return (double)(SP(buf, q, pa) - SP(buf, q, pb));
    };
    auto DKD = [&](int bu, int bc, int q, int pp) -> double {
        if (ANOM)
            return ((double)SP(bu, q, pp) - (double)SP(bc, q, pp)) + ((double)SR(bu, q, pp) - (double)SR(bc, q, pp));
        return (double)(SP(bu, q, pp) - SP(bc, q, pp));
    };
    // the isoneutral slope terms of a cell at level k < maxk from the staged
    // T and S (2489-2501, 2561-2605): dzrho, the four corner gradients, the
    // slope limiter; counts limps and the diffusive stability number
    auto iso_coeffs = [&](int k, int cb, int ub, bool& iso, double& dzrho, double& rdzrho, double& slim,
                          double* dxr, double* dyr) {
        const AS_GLOBAL double* mkD = gp(g.mk);
        const double rdza = mkD[MK_RDZA * KS + k], dt = mkD[MK_DT * KS + k];
        const double Tc = SPD(cb, qT, p), Tu = SPD(ub, qT, p);
        const double tatw = 0.5 * (Tc + Tu);
        const double tec = -e0D - e2D * tatw * 2 - e3D * tatw * tatw * 3;
        dzrho = (e1D * DKD(ub, cb, qS, p) - tec * DKD(ub, cb, qT, p)) * rdza;
        if (M.diso && dzrho < -1.0E-12) {
            iso = true;
            rdzrho = fdiv(1.0, dzrho);
            double tv1 = 0;
            // ina order (knp, nnp) = (0,0), (0,1), (1,0), (1,1)
#pragma unroll
            for (int n = 0; n < 4; n++) {
                const int b = n < 2 ? cb : ub;
                const int kn = n < 2 ? k : k + 1;
                const bool east = n & 1;
                const bool mx = kn >= (east ? k1e : k1w), my = kn >= (east ? k1n : k1s);
                const double dxT = mx ? (east ? DD(b, qT, pe, p) : DD(b, qT, p, pw)) * rcD * rdphiD : 0.0;
                const double dxS = mx ? (east ? DD(b, qS, pe, p) : DD(b, qS, p, pw)) * rcD * rdphiD : 0.0;
                const double dyT = my ? (east ? DD(b, qT, pn, p) * cvnD * rdsvnD : DD(b, qT, p, ps) * cvsD * rdsvsD) : 0.0;
                const double dyS = my ? (east ? DD(b, qS, pn, p) * cvnD * rdsvnD : DD(b, qS, p, ps) * cvsD * rdsvsD) : 0.0;
                dxr[n] = sccD * dxS - tec * dxT;
                dyr[n] = sccD * dyS - tec * dyT;
                tv1 = tv1 + dxr[n] * dxr[n] + dyr[n] * dyr[n];
            }
            tv1 = 0.25 * tv1 * rdzrho * rdzrho;
            const double ssmax = mkD[MK_SSMAX * KS + k];
            if (tv1 > ssmax) {
                slim = fdiv(ssmax * ssmax, tv1 * tv1);
                limps++;
            } else {
                slim = 1;
            }
            tv1 = tv1 * slim * diff1D * rdza;
            const double tvd = tv1 * dt * rdza;
            if (tvd > dmaxv) dmaxv = tvd;
        }
    };
    // the FP32 form's isoneutral coefficients: corner differences times ax, ay
    // (the sign of the west/south corners folded in) and the vertical qq,
    // formed in FP64 and rounded once
    auto fp32_iso = [&](int k, bool iso, double slim, double rdzrho, const double* dxr, const double* dyr, R* ax,
                        R* ay, R& qq) {
#pragma unroll
        for (int n = 0; n < 4; n++) ax[n] = ay[n] = (R)0;
        qq = (R)0;
        if (iso) {
            const double rdza = gp(g.mk)[MK_RDZA * KS + k];
            const double isoc = 0.25 * slim * diff1D;
            const double f = 2 * isoc * rdzrho;
            double s2 = 0;
#pragma unroll
            for (int n = 0; n < 4; n++) {
                const bool east = n & 1;
                ax[n] = (R)(f * dxr[n] * rcD * rdphiD * (east ? 1.0 : -1.0));
                ay[n] = (R)(f * dyr[n] * (east ? cvnD * rdsvnD : -(cvsD * rdsvsD)));
                s2 = s2 + dxr[n] * dxr[n] + dyr[n] * dyr[n];
            }
            qq = (R)(isoc * s2 * (rdzrho * rdzrho) * rdza);
        }
    };
    // Members with BIOGEM: the coupling's column sums of the new state
    // (biogem.f90:1984-1997, 2016-2025), accumulated up the column in the
    // order the coupling adds them; co redoes the columns it mixes.
    const bool sums = IM != 1 && M.bgc != 0;
    const double rSold = sums ? 1.0 / gp(M.scal)[SC_SMEAN] : 0.0;
    double accs[LB], accS = 0.0;
#pragma unroll
    for (int q = 0; q < LB; q++) accs[q] = 0.0;
    auto add_sum = [&](int q, double nv, double os, double mm, double vv) {
        const int l = l0 + q;
        if (l == 1) {
            accS += (nv + SALN0 + 0.0) * vv;
        } else if (l >= 2) {
            const double lv = nv * os * rSold;
            accs[q] += lv * mm;
        }
    };
    fetch(kmin - 1);
    commit(kmin % FLUX_NBUF);
    if (kmin < maxk) {
        fetch(kmin);
        commit((kmin + 1) % FLUX_NBUF);
    }
    __syncthreads();
    for (int k = kmin; k <= maxk; k++) {
        const int kz = k - 1;
        const int cb = k % FLUX_NBUF, ub = (k + 1) % FLUX_NBUF;
        const bool top = (k == maxk);
        const bool more = k + 2 <= maxk;
        if (more) fetch(kz + 2);
        if (IM == 1 && active && k >= k1c && !top && M.diso) {
            // the pre-pass: the cell's isoneutral coefficients, once
            const uint32_t oc = (uint32_t)(kz * plane + col);
            bool iso = false;
            double dzrho = 0, rdzrho = 0, slim = 1, dxr[4] = {0, 0, 0, 0}, dyr[4] = {0, 0, 0, 0};
            iso_coeffs(k, cb, ub, iso, dzrho, rdzrho, slim, dxr, dyr);
            if constexpr (ANOM) {
                R ax[4], ay[4], qq;
                fp32_iso(k, iso, slim, rdzrho, dxr, dyr, ax, ay, qq);
                AS_GLOBAL float* ic = reinterpret_cast<AS_GLOBAL float*>(gp(M.isoc));
#pragma unroll
                for (int n = 0; n < 4; n++) {
                    at(ic, (uint32_t)(n * ncell) + oc) = ax[n];
                    at(ic, (uint32_t)((4 + n) * ncell) + oc) = ay[n];
                }
                at(ic, (uint32_t)(8 * ncell) + oc) = qq;
            } else {
                AS_GLOBAL double* ic = gp(M.isoc);
                at(ic, oc) = iso ? (double)dzrho : 1.0;  // +1: no isoneutral flux in this cell
#pragma unroll
                for (int n = 0; n < 4; n++) {
                    at(ic, (uint32_t)((1 + n) * ncell) + oc) = dxr[n];
                    at(ic, (uint32_t)((5 + n) * ncell) + oc) = dyr[n];
                }
                at(ic, (uint32_t)(9 * ncell) + oc) = slim;
            }
        }
        if (IM != 1 && active && k >= k1c) {
            const int o = kz * plane;
            // face velocities and Peclet-weighted upwinding (goldstein.f90:2489-2523);
            // the FP32 form reads its face coefficients instead (below)
            const uint32_t oc = (uint32_t)(o + col);
            const R dza = mk[MK_DZA * KS + k], rdza = mk[MK_RDZA * KS + k];
            const R dt = mk[MK_DT * KS + k], rdz = mk[MK_RDZ * KS + k];
            R ue_ = 0, uw_ = 0, un_ = 0, us_ = 0, ua_ = 0;
            R upe = 0, upw = 0, upn = 0, ups = 0, upa = 0;
            if constexpr (!ANOM) {
                ue_ = at(u, oc);
                uw_ = at(u, (uint32_t)(o + cw));
                un_ = at(u, ncell + oc);
                us_ = at(u, (uint32_t)(ncell + o + cs));
                ua_ = at(u, 2 * ncell + oc);
                R pec;
                pec = ue_ * dphi / diff1; upe = pec / ((R)2.0 + fabs(pec));
                pec = uw_ * dphi / diff1; upw = pec / ((R)2.0 + fabs(pec));
                pec = un_ * dsvn / diff1; upn = pec / ((R)2.0 + fabs(pec));
                pec = us_ * dsvs / diff1; ups = pec / ((R)2.0 + fabs(pec));
                pec = ua_ * dza / diffv; upa = pec / ((R)2.0 + fabs(pec));
            }
            const bool oe = k >= max(k1c, k1e), ow = k >= max(k1w, k1c);
            const bool on = k >= max(k1c, k1n), os = k >= max(k1s, k1c);
            // isoneutral coefficients from T and S (2489-2501, 2561-2605)
            bool iso = false;
            double dzrho = 0, rdzrho = 0, slim = 1, dxr[4] = {0, 0, 0, 0}, dyr[4] = {0, 0, 0, 0};
            R ax[4] = {0, 0, 0, 0}, ay[4] = {0, 0, 0, 0}, qq = 0;  // the FP32 form's coefficients
            if (!top) {
                if constexpr (IM == 2) {
                    if (M.diso) {
                        if constexpr (ANOM) {
                            // zero coefficients give the flux without the isoneutral term exactly
                            const AS_GLOBAL float* ic = reinterpret_cast<const AS_GLOBAL float*>(gp(M.isoc));
#pragma unroll
                            for (int n = 0; n < 4; n++) {
                                ax[n] = at(ic, (uint32_t)(n * ncell) + oc);
                                ay[n] = at(ic, (uint32_t)((4 + n) * ncell) + oc);
                            }
                            qq = at(ic, (uint32_t)(8 * ncell) + oc);
                            iso = true;
                        } else {
                            const AS_GLOBAL double* ic = gp(M.isoc);
                            dzrho = at(ic, oc);
                            iso = dzrho < 0.0;
                            if (iso) {
#pragma unroll
                                for (int n = 0; n < 4; n++) {
                                    dxr[n] = at(ic, (uint32_t)((1 + n) * ncell) + oc);
                                    dyr[n] = at(ic, (uint32_t)((5 + n) * ncell) + oc);
                                }
                                slim = at(ic, (uint32_t)(9 * ncell) + oc);
                            }
                        }
                    }
                } else {
                    iso_coeffs(k, cb, ub, iso, dzrho, rdzrho, slim, dxr, dyr);
                    if constexpr (ANOM) fp32_iso(k, iso, slim, rdzrho, dxr, dyr, ax, ay, qq);
                }
            }
            // discrete divergence of the open faces (FP32 form only)
            // (evaluated in FP64 from the FP64 velocities when they are set)
            const R divu = ANOM ? at(u, 6 * ncell + oc) : (R)0;
            const R isoc = (R)0.25 * slim * diff1;
            double newT = 0, newS = 0;
            double s_os = 0.0, s_mm = 0.0, s_vv = 0.0;
            if (sums) {
                s_os = at(gp(M.ocn), ncell + oc);
                s_mm = at(gp(M.M), oc);
                if (l0 == 0) s_vv = at(gp(g.V), oc);
            }
            if constexpr (ANOM) {
                // FP32 form in coefficient form: each face flux is the
                // neighbour difference times a coefficient of the cell, and the
                // isoneutral flux a sum of corner differences times
                // coefficients.  The face coefficients depend on the velocity,
                // the diffusivities and the mask only: cgo_set_velocity forms
                // them in FP64 and rounds each once (closed faces are zero)
                const R ce = at(u, oc), cw = at(u, ncell + oc);
                const R cn = at(u, 2 * ncell + oc), cs = at(u, 3 * ncell + oc);
                const R ca = at(u, 4 * ncell + oc), cbn = at(u, 5 * ncell + oc);
#pragma unroll
                for (int q = 0; q < LB; q++) {
                    if (q >= nl) break;
                    const R Tc = SP(cb, q, p);
                    const R dE = D(cb, q, pe, p), dW = D(cb, q, pw, p), dN = D(cb, q, pn, p), dS = D(cb, q, ps, p);
                    R fa, fbn = 0;
                    // the top cell's surface flux enters in FP64: rounded to
                    // float it would bias every step's forcing alike
                    double fsurf = 0.0;
                    if (top) {
                        fa = (R)0;
                        if (M.sflux) fsurf = at(gp(M.sflux), (uint32_t)((l0 + q) * plane + col));
                    } else {
                        const R dU = DK(ub, cb, q, p);
                        R tv = 0;
                        if (iso) {
                            // corner differences with the sign of the west/south
                            // corners folded into ax[0], ax[2], ay[0], ay[2]
                            tv = ax[0] * dW + ax[1] * dE + ay[0] * dS + ay[1] * dN;
                            tv = tv + ax[2] * D(ub, q, pw, p) + ax[3] * D(ub, q, pe, p) + ay[2] * D(ub, q, ps, p) +
                                 ay[3] * D(ub, q, pn, p);
                            tv = tv - qq * dU;
                        }
                        fa = dU * ca + tv;
                        fbn = dU * cbn + tv;
                    }
                    const R div = ((dE * ce - dW * cw) + (dN * cn - dS * cs)) + (fa - FB(q)) * rdz;
                    const int idx = (l0 + q) * tstride + o + col;
                    const R inc = -(dt * (div + Tc * divu));
                    const double incs = fsurf != 0.0 ? -(gp(g.mk)[MK_DT * KS + k] * (fsurf * gp(g.mk)[MK_RDZ * KS + k])) : 0.0;
                    double nv;
                    if (lo_src) {
                        // compensated accumulation: the rounding residual of every
                        // update is carried to the next step in the int16 plane
                        nv = ((double)Tc + (double)SR(cb, q, p)) + (double)inc + incs;
                        StateIO<R>::st((R*)dst, lo_dst, idx, nv);
                    } else {
                        const R v = Tc + (inc + (R)incs);
                        at(dst, (uint32_t)idx) = v;
                        nv = (double)v;
                    }
                    FB(q) = fbn;
                    if (l0 + q == 0) newT = nv;
                    if (l0 + q == 1) newS = nv;
                    if (sums) add_sum(q, nv, s_os, s_mm, s_vv);
                }
            } else
#pragma unroll
            for (int q = 0; q < LB; q++) {
                if (q >= nl) break;
                const R Tc = SP(cb, q, p), Te = SP(cb, q, pe), Tw = SP(cb, q, pw), Tn = SP(cb, q, pn), Ts = SP(cb, q, ps);
                R fe, fw, fn, fs, fa, fbn = 0;
                if (!ANOM) {
                    fe = oe ? ue_ * rc * (((R)1.0 - upe) * Te + ((R)1.0 + upe) * Tc) * (R)0.5 - (Te - Tc) * rc2 * diff1 : (R)0;
                    fw = ow ? uw_ * rc * (((R)1.0 - upw) * Tc + ((R)1.0 + upw) * Tw) * (R)0.5 - (Tc - Tw) * rc2 * diff1 : (R)0;
                    fn = on ? cvn * un_ * (((R)1.0 - upn) * Tn + ((R)1.0 + upn) * Tc) * (R)0.5 - cv2n * (Tn - Tc) * diff1 : (R)0;
                    fs = os ? cvs * us_ * (((R)1.0 - ups) * Tc + ((R)1.0 + ups) * Ts) * (R)0.5 - cv2s * (Tc - Ts) * diff1 : (R)0;
                } else {
                    const R dE = D(cb, q, pe, p), dW = D(cb, q, pw, p), dN = D(cb, q, pn, p), dS = D(cb, q, ps, p);
                    fe = oe ? ue_ * rc * ((R)1.0 - upe) * dE * (R)0.5 - dE * rc2 * diff1 : (R)0;
                    fw = ow ? uw_ * rc * ((R)1.0 + upw) * dW * (R)0.5 + dW * rc2 * diff1 : (R)0;
                    fn = on ? cvn * un_ * ((R)1.0 - upn) * dN * (R)0.5 - cv2n * dN * diff1 : (R)0;
                    fs = os ? cvs * us_ * ((R)1.0 + ups) * dS * (R)0.5 + cv2s * dS * diff1 : (R)0;
                }
                if (top) {
                    fa = M.sflux ? (R)at(gp(M.sflux), (uint32_t)((l0 + q) * plane + col)) : (R)0;
                } else {
                    const R Tu = SP(ub, q, p);
                    const R dUc = DK(ub, cb, q, p);
                    R nadv = -dUc * rdza * diffv;
                    if (!ANOM) {
                        fa = ua_ * (((R)1.0 - upa) * Tu + ((R)1.0 + upa) * Tc) * (R)0.5;
                        fa = fa - (Tu - Tc) * rdza * diffv;
                    }
                    if (iso) {
                        const R dzts = dUc * rdza;
                        R tv = 0;
#pragma unroll
                        for (int n = 0; n < 4; n++) {
                            const int b = n < 2 ? cb : ub;
                            const bool east = n & 1;
                            const int kn = n < 2 ? k : k + 1;
                            const R dx = kn >= (east ? k1e : k1w)
                                             ? (east ? D(b, q, pe, p) : D(b, q, p, pw)) * rc * rdphi
                                             : (R)0;
                            const R dy = kn >= (east ? k1n : k1s)
                                             ? (east ? D(b, q, pn, p) * cvn * rdsvn : D(b, q, p, ps) * cvs * rdsvs)
                                             : (R)0;
                            tv = tv + ((R)2 * dzrho * dx - dxr[n] * dzts) * dxr[n] + ((R)2 * dzrho * dy - dyr[n] * dzts) * dyr[n];
                        }
                        tv = isoc * tv / (dzrho * dzrho);
                        if (!ANOM) fa = fa + tv;
                        nadv = nadv + tv;
                    }
                    if (ANOM) {
                        fa = ua_ * ((R)1.0 - upa) * dUc * (R)0.5 + nadv;       // relative to this cell
                        fbn = -(ua_ * ((R)1.0 + upa) * dUc * (R)0.5) + nadv;   // relative to the cell above
                    }
                }
                const R div = ((fe - fw) * rdphi + (fn - fs) * rds) + (fa - FB(q)) * rdz;
                const int idx = (l0 + q) * tstride + o + col;
                double nv;
                if (ANOM) {
                    const R inc = -(dt * (div + Tc * divu));
                    if (lo_src) {
                        // compensated accumulation: the rounding residual of every
                        // update is carried to the next step in the int16 plane
                        nv = ((double)Tc + (double)SR(cb, q, p)) + (double)inc;
                        StateIO<R>::st((R*)dst, lo_dst, idx, nv);
                    } else {
                        const R v = Tc + inc;
                        at(dst, (uint32_t)idx) = v;
                        nv = (double)v;
                    }
                    FB(q) = fbn;
                } else {
                    const R v = Tc - dt * div;
                    at(dst, (uint32_t)idx) = v;
                    nv = (double)v;
                    FB(q) = fa;
                }
                if (l0 + q == 0) newT = nv;
                if (l0 + q == 1) newS = nv;
                if (sums) add_sum(q, nv, s_os, s_mm, s_vv);
            }
            if (l0 == 0) at(rhop, oc) = eos<double>(g.ec[0], g.ec[1], g.ec[2], g.ec[3], newT, newS);
        }
        if (FLUX_NBUF == 2) __syncthreads();  // plane k is read by every thread before it is replaced
        if (more) commit((k + 2) % FLUX_NBUF);
        __syncthreads();
    }
    if (IM != 2 && l0 == 0 && M.diso && active) {
        if (limps) atomicAdd(&M.diag[0], limps);
        if (dmaxv > 0.0) atomic_max_pos(M.dmax, dmaxv);
    }
    if (sums && active && k1c <= maxk) {
        const int slot = gp(g.cslot)[col];
        AS_GLOBAL double* cs = gp(M.colsum);
#pragma unroll
        for (int q = 0; q < LB; q++) {
            const int l = l0 + q;
            if (q >= nl) break;
            if (l == 1) cs[slot] = accS;
            else if (l >= 2) cs[(size_t)l * g.nwet + slot] = accs[q];
        }
    }
}

// Launch of k_flux over a batch: nt threads per block (128 or 256), lb
// tracers per block, nbuf level planes; with iso_split the isoneutral
// coefficients come from a pre-pass launch (IM 1) and the tracer blocks (IM 2)
// stage only their tracers.  part: 1 the pre-pass, 2 the tracer blocks, 3 both.  The tracer block, then the block width, then the
// ring shrink until the workgroup fits the LDS the device grants one block.
template <typename R>
void launch_flux_t(const MemberDev* t, const int32_t* l, int n, const FluxShape& fs, int maxL, int lb, int nt,
                   int nbuf, int iso_split, int part, size_t lds_cap, hipStream_t s) {
    constexpr int es = sizeof(R);
    if (nt != 256) nt = 128;
    if (nbuf != 2) nbuf = 3;
    auto bytes = [&](int ns, int b) { return flux_lds_bytes(ns, b, es, nt, nbuf); };
    auto slots = [&]() { return iso_split ? lb : lb + 2; };
    while (lb > 2 && bytes(slots(), lb) > lds_cap) lb = lb == 8 ? 6 : (lb == 6 ? 4 : 2);
    if (nt == 256 && bytes(slots(), lb) > lds_cap) nt = 128;
    if (nbuf == 3 && bytes(slots(), lb) > lds_cap) nbuf = 2;
    const int v = nt == 256;
    if (fs.nblk[v] == 0) return;
    const dim3 grid(fs.nblk[v], (maxL + lb - 1) / lb, n);
    const size_t lds = bytes(slots(), lb);
#define CGO_FLUX(LB_, NT_, NB_, IM_) \
    hipLaunchKernelGGL((k_flux<R, LB_, NT_, NB_, IM_>), grid, dim3(NT_), lds, s, t, l)
#define CGO_FLUX_LB(NT_, NB_, IM_)                    \
    do {                                              \
        if (lb == 2) CGO_FLUX(2, NT_, NB_, IM_);      \
        else if (lb == 8) CGO_FLUX(8, NT_, NB_, IM_); \
        else if (lb == 6) CGO_FLUX(6, NT_, NB_, IM_); \
        else CGO_FLUX(4, NT_, NB_, IM_);              \
    } while (0)
    if (iso_split) {
        // the pre-pass: one block per column block, T and S staged
        const dim3 g1(fs.nblk[v], 1, n);
        if (part & 1) {
        const size_t lds1 = bytes(2, 2);
        if (nt == 256) {
            if (nbuf == 2) hipLaunchKernelGGL((k_flux<R, 2, 256, 2, 1>), g1, dim3(256), lds1, s, t, l);
            else hipLaunchKernelGGL((k_flux<R, 2, 256, 3, 1>), g1, dim3(256), lds1, s, t, l);
        } else {
            if (nbuf == 2) hipLaunchKernelGGL((k_flux<R, 2, 128, 2, 1>), g1, dim3(128), lds1, s, t, l);
            else hipLaunchKernelGGL((k_flux<R, 2, 128, 3, 1>), g1, dim3(128), lds1, s, t, l);
        }
        }
        if (!(part & 2)) {
        } else if (nt == 256) {
            if (nbuf == 2) CGO_FLUX_LB(256, 2, 2); else CGO_FLUX_LB(256, 3, 2);
        } else {
            if (nbuf == 2) CGO_FLUX_LB(128, 2, 2); else CGO_FLUX_LB(128, 3, 2);
        }
    } else if (!(part & 2)) {
        // no pre-pass: the tracer blocks compute their own coefficients
    } else if (nt == 256) {
        if (nbuf == 2) CGO_FLUX_LB(256, 2, 0); else CGO_FLUX_LB(256, 3, 0);
    } else {
        if (nbuf == 2) CGO_FLUX_LB(128, 2, 0); else CGO_FLUX_LB(128, 3, 0);
    }
#undef CGO_FLUX_LB
#undef CGO_FLUX
}

// co, goldstein.f90:2657-2777 (iconv = 0, ieos = 0): FP64 arithmetic on the
// new state in place, in two passes over a workgroup of CB <= blockDim.x wet
// columns.  Pass 1, a thread per column, all its density loads issued
// together: a column whose density falls strictly from k1 up -- most
// columns; the reference loop leaves them as they are -- only has its top T
// and S noted; the others are listed in LDS.  Pass 2: each group of CO_LANES
// lanes takes the next listed column whenever it has finished one.  The
// column's level arrays (segment map, thicknesses, densities and its T and S
// as the state stores them) sit in LDS, one copy per column: its lanes walk
// the reference's mixing loop in lockstep and write the same values, so they
// reach the same decisions; each lane mixes and copies its own share of the
// tracers in memory (l = lane, lane + CO_LANES, ...), each tracer's
// operations in the reference's order.  K = the level arrays' length (the
// launch's deepest maxk + 2).  Fewer columns per workgroup (CB) make more
// workgroups, which hold more convecting columns per CU at once where most
// convect, and more that only run pass 1 where few do.
constexpr int CO_LANES = 4;
constexpr int CO_CH = 4;  // tracers whose loads one lane issues together
__host__ __device__ constexpr size_t co_lds_bytes(int nt, int cb, int K) {
    return (size_t)(nt / CO_LANES) * K * (4 * 8 + 4) + (size_t)cb * 4;
}
template <typename R>
__global__ __launch_bounds__(256) void k_co(const MemberDev* __restrict__ mems, const int32_t* __restrict__ list, int K, int CB) {
    extern __shared__ double co_lds[];
    __shared__ int s_n, s_next;
    const BlockXYZ B = xcd_blk();
    const MemberDev& M = mems[list[B.y]];
    const GridDev& g = M.g;
    const int nt = (int)blockDim.x, P = nt / CO_LANES;
    const int maxi = g.maxi, maxk = g.maxk, L = M.L;
    const int par = *gp(M.parity);
    void* st = M.st[par ^ 1];
    int16_t* lo = M.lo[par ^ 1];
    const bool carry = lo != nullptr;
    const size_t plane = (size_t)g.ncol, tstride = (size_t)maxk * plane;
    AS_GLOBAL double* rho = gp(M.rho);
    int* wl = reinterpret_cast<int*>(co_lds + (size_t)P * 4 * K) + (size_t)P * K;  // the list of pass 1
    // members with surface fluxes keep the top level's T and S as tstepo
    // leaves them (tsval_ocn, ssval_ocn of step_goldstein 427-430): the
    // next call of the fluxes reads them, whatever the coupling does after
    auto snap = [&](int col, double t, double s) {
        if (M.sw) {
            gp(M.sw)[(size_t)SW_TSTAR * plane + col] = t;
            gp(M.sw)[(size_t)SW_SSTAR * plane + col] = s;
        }
    };
    if (threadIdx.x == 0) s_n = 0, s_next = 0;
    __syncthreads();
    // ---- pass 1: the loop below walks a stable column's m down to k1
    // without mixing and the fill-in finds every level in place
    {
        const int w = B.x * CB + (int)threadIdx.x;
        if ((int)threadIdx.x < CB && w < g.nwet) {
            const int col = g.wetsorted[w];
            const int k1 = k1at(g, col % maxi + 1, col / maxi + 1);
            bool stable = true;
            double rb = rho[(size_t)(maxk - 1) * plane + col];
            for (int q = maxk - 1; q >= k1; q--) {
                const double ra = rho[(size_t)(q - 1) * plane + col];
                stable = stable & (rb < ra);
                rb = ra;
            }
            if (stable) {
                const size_t top = (size_t)(maxk - 1) * plane + col;
                snap(col, StateIO<R>::ld(st, lo, top), StateIO<R>::ld(st, lo, tstride + top));
            } else {
                wl[atomicAdd(&s_n, 1)] = w;
            }
        }
    }
    __syncthreads();
    const int n = s_n;
    if (n == 0) return;  // uniform for the block
    // ---- pass 2
    const int sub = (int)threadIdx.x % CO_LANES, cl = (int)threadIdx.x / CO_LANES;
    const AS_GLOBAL double* dz = gp(g.mk) + MK_DZ * (maxk + 2);
    const double e0 = g.ec[0], e1 = g.ec[1], e2 = g.ec[2], e3 = g.ec[3];
    double* r = co_lds + (size_t)cl * 4 * K;
    double* dzm = r + K;
    double* tp = dzm + K;
    double* sp = tp + K;
    int* kk = reinterpret_cast<int*>(co_lds + (size_t)P * 4 * K) + (size_t)cl * K;
    // this lane's tracers: l = sub + CO_LANES * j, j < nmine
    const int nmine = L > sub ? (L - sub + CO_LANES - 1) / CO_LANES : 0;
    // the group's lanes sit side by side in one wave; its first lane takes
    // the next list entry for the group
    const int lead = ((int)threadIdx.x & 63) & ~(CO_LANES - 1);
    for (;;) {
        int idx = 0;
        if (sub == 0) idx = atomicAdd(&s_next, 1);
        idx = __shfl(idx, lead);
        if (idx >= n) break;
        {
            const int w = wl[idx];
            const int col = g.wetsorted[w];
            const int k1 = k1at(g, col % maxi + 1, col / maxi + 1);
            for (int m = 0; m <= maxk + 1; m++) kk[m] = 0;
            for (int m = k1; m <= maxk; m++) {
                r[m] = rho[(size_t)(m - 1) * plane + col];
                tp[m] = StateIO<R>::ld(st, lo, (size_t)(m - 1) * plane + col);
                sp[m] = StateIO<R>::ld(st, lo, tstride + (size_t)(m - 1) * plane + col);
                kk[m] = m;
                dzm[m] = dz[m];
            }
            int m = maxk, lastmix = 0;
            bool mixed = false;
            // every pass either moves m or merges two segments, so the
            // reference loop ends within O(maxk^2) passes; the cap only
            // guards non-finite densities
            int guard = 4 * maxk * maxk + 64;
            while ((kk[m - 1] > 0 || (lastmix != 0 && kk[m] != maxk)) && --guard > 0) {
                if (kk[m - 1] == 0 || r[kk[m]] < r[kk[m - 1]]) {
                    if (lastmix == 0 || kk[m] == maxk)
                        m = m - 1;
                    else
                        m = m + 1;
                    lastmix = 0;
                } else {
                    lastmix = 1;
                    mixed = true;
                    int nn = m - 1;
                    while (kk[nn - 1] > 0 && r[kk[nn]] >= r[kk[nn - 1]]) nn = nn - 1;
                    const int km = kk[m];
                    double dsum = dzm[km];
                    for (int ni = 1; ni <= m - nn; ni++) dsum = dsum + dzm[kk[m - ni]];
                    // the thickness-weighted mean of levels kk[nn..m]: T and S
                    // privately on every lane, this lane's tracers in memory,
                    // CO_CH at a time (their loads issued together)
                    {
                        double at = tp[km] * dzm[km], as = sp[km] * dzm[km];
                        for (int ni = 1; ni <= m - nn; ni++) {
                            const int kq = kk[m - ni];
                            at = at + tp[kq] * dzm[kq];
                            as = as + sp[kq] * dzm[kq];
                        }
                        tp[km] = StateIO<R>::rt(at / dsum, carry);
                        sp[km] = StateIO<R>::rt(as / dsum, carry);
                    }
                    for (int j0 = 0; j0 < nmine; j0 += CO_CH) {
                        double sm[CO_CH];
#pragma unroll
                        for (int c = 0; c < CO_CH; c++)
                            if (j0 + c < nmine)
                                sm[c] = StateIO<R>::ld(st, lo, (size_t)(sub + CO_LANES * (j0 + c)) * tstride + col +
                                                                   (size_t)(km - 1) * plane) * dzm[km];
                        for (int ni = 1; ni <= m - nn; ni++) {
                            const int kq = kk[m - ni];
                            const double wq = dzm[kq];
#pragma unroll
                            for (int c = 0; c < CO_CH; c++)
                                if (j0 + c < nmine)
                                    sm[c] = sm[c] + StateIO<R>::ld(st, lo, (size_t)(sub + CO_LANES * (j0 + c)) * tstride +
                                                                               col + (size_t)(kq - 1) * plane) * wq;
                        }
#pragma unroll
                        for (int c = 0; c < CO_CH; c++)
                            if (j0 + c < nmine)
                                StateIO<R>::st(st, lo, (size_t)(sub + CO_LANES * (j0 + c)) * tstride + col + (size_t)(km - 1) * plane,
                                               sm[c] / dsum);
                    }
                    for (int ni = 1; ni <= m - nn; ni++) dzm[km] = dzm[km] + dzm[kk[m - ni]];
                    r[km] = eos<double>(e0, e1, e2, e3, tp[km], sp[km]);
                    int ni = m - 1;
                    while (kk[ni + 1] > 0) {
                        kk[ni] = kk[ni - m + nn];
                        ni = ni - 1;
                    }
                }
            }
            if (guard <= 0 && sub == 0) atomicAdd(&M.diag[3], 1);
            // fill-in: every level not left as a segment's own takes the
            // values of the segment it lies in
            m = maxk - 1;
            double cost = 0.0;
            for (int nn = maxk - 1; nn >= k1; nn--) {
                if (nn > kk[m]) {
                    const int kf = kk[m + 1];
                    const size_t from = (size_t)(kf - 1) * plane + col, to = (size_t)(nn - 1) * plane + col;
                    for (int j0 = 0; j0 < nmine; j0 += CO_CH) {
                        double v[CO_CH];
#pragma unroll
                        for (int c = 0; c < CO_CH; c++)
                            if (j0 + c < nmine) v[c] = StateIO<R>::ld(st, lo, (size_t)(sub + CO_LANES * (j0 + c)) * tstride + from);
#pragma unroll
                        for (int c = 0; c < CO_CH; c++)
                            if (j0 + c < nmine) StateIO<R>::st(st, lo, (size_t)(sub + CO_LANES * (j0 + c)) * tstride + to, v[c]);
                    }
                    tp[nn] = tp[kf];
                    sp[nn] = sp[kf];
                    r[nn] = eos<double>(e0, e1, e2, e3, tp[nn], sp[nn]);
                    cost += 1.0;
                } else {
                    m = m - 1;
                }
            }
            if ((mixed || cost > 0.0) && M.bgc) {
                // the coupling's column sums of this column after mixing and fill-in
                const double rSold = 1.0 / gp(M.scal)[SC_SMEAN];
                const size_t ncl = (size_t)g.ncell, slot = (size_t)gp(g.vpos)[w];
                AS_GLOBAL double* cs = gp(M.colsum);
                if (sub == 0) {
                    double aS = 0.0;
                    for (int q = k1; q <= maxk; q++) aS += (sp[q] + SALN0 + 0.0) * gp(g.V)[(size_t)(q - 1) * plane + col];
                    cs[slot] = aS;
                }
                // this lane's tracers from the third on
                const int jfirst = sub >= 2 ? 0 : 1;
                for (int j0 = jfirst; j0 < nmine; j0 += CO_CH) {
                    double a[CO_CH];
#pragma unroll
                    for (int t = 0; t < CO_CH; t++) a[t] = 0.0;
                    for (int q = k1; q <= maxk; q++) {
                        const size_t c = (size_t)(q - 1) * plane + col;
                        const double so = gp(M.ocn)[ncl + c], mm = gp(M.M)[c];
#pragma unroll
                        for (int t = 0; t < CO_CH; t++)
                            if (j0 + t < nmine) {
                                const double lv =
                                    StateIO<R>::ld(st, lo, (size_t)(sub + CO_LANES * (j0 + t)) * tstride + c) * so * rSold;
                                a[t] += lv * mm;
                            }
                    }
#pragma unroll
                    for (int t = 0; t < CO_CH; t++)
                        if (j0 + t < nmine) cs[(size_t)(sub + CO_LANES * (j0 + t)) * g.nwet + slot] = a[t];
                }
            }
            if (sub == 0) {
                if (mixed || cost > 0.0) {
                    for (int q = k1; q <= maxk; q++) rho[(size_t)(q - 1) * plane + col] = r[q];
                    if (cost > 0.0) gp(M.cost)[col] += cost;
                }
                snap(col, tp[maxk], sp[maxk]);
            }
        }
    }
}

#ifdef CGO_DEFINE_FLIP
// end of the GOLDSTEIN step: flip the ping-pong parity, count the step
__global__ void k_flip(const MemberDev* __restrict__ mems, const int32_t* __restrict__ list, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const MemberDev& M = mems[list[t]];
    *M.parity ^= 1;
    *M.steps += 1;
}
#endif

}  // namespace cgo
