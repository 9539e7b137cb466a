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

constexpr int FLUX_THREADS = 128;

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
__device__ __forceinline__ double carry_unit(float x) {
    int e = (x == 0.0f) ? -126 : ilogbf(x);
    if (e < -126) e = -126;
    return ldexp(1.0, e - 23 - 16);
}

// state access as double (FP32 form: hi + carry)
template <typename R> struct StateIO;
template <> struct StateIO<double> {
    static __device__ __forceinline__ double ld(const void* s, const int16_t*, size_t i) {
        return static_cast<const double*>(s)[i];
    }
    static __device__ __forceinline__ void st(void* s, int16_t*, size_t i, double x) {
        static_cast<double*>(s)[i] = x;
    }
};
template <> struct StateIO<float> {
    static __device__ __forceinline__ double ld(const void* s, const int16_t* lo, size_t i) {
        float h = static_cast<const float*>(s)[i];
        double x = (double)h;
        if (lo) x += (double)lo[i] * carry_unit(h);
        return x;
    }
    static __device__ __forceinline__ void st(void* s, int16_t* lo, size_t i, double x) {
        float h = (float)x;
        static_cast<float*>(s)[i] = h;
        if (lo) {
            double r = rint((x - (double)h) / carry_unit(h));
            r = fmin(fmax(r, -32767.0), 32767.0);
            lo[i] = (int16_t)r;
        }
    }
};

__device__ __forceinline__ void atomic_max_pos(double* p, double v) {
    // non-negative doubles order like their bit patterns
    atomicMax(reinterpret_cast<unsigned long long*>(p), (unsigned long long)__double_as_longlong(v));
}

template <typename R, int LB>
__global__ __launch_bounds__(FLUX_THREADS) void k_flux(const MemberDev* __restrict__ mems,
                                                       const int32_t* __restrict__ list) {
    const MemberDev& M = mems[list[blockIdx.z]];
    const GridDev& g = M.g;
    const int maxi = g.maxi, maxj = g.maxj, maxk = g.maxk, L = M.L;
    const int col = blockIdx.x * FLUX_THREADS + threadIdx.x;
    const int l0 = blockIdx.y * LB;
    if (col >= g.ncol || l0 >= L) return;
    const int i0 = col % maxi, j0 = col / maxi;
    const int I = i0 + 1, J = j0 + 1;
    const int k1c = k1at(g, I, J);
    if (k1c > maxk) return;  // land column: nothing is updated
    const int k1e = k1at(g, I + 1, J), k1w = k1at(g, I - 1, J);
    const int k1n = k1at(g, I, J + 1), k1s = k1at(g, I, J - 1);
    const int par = *M.parity;
    const R* __restrict__ src = static_cast<const R*>(M.st[par]);
    R* __restrict__ dst = static_cast<R*>(M.st[par ^ 1]);
    const int16_t* __restrict__ lo_src = M.lo[par];
    int16_t* __restrict__ lo_dst = M.lo[par ^ 1];
    const R* __restrict__ u = static_cast<const R*>(M.u);
    const size_t plane = (size_t)g.ncol, tstride = (size_t)maxk * plane;
    const int ce = j0 * maxi + (i0 + 1 == maxi ? 0 : i0 + 1);
    const int cw = j0 * maxi + (i0 == 0 ? maxi - 1 : i0 - 1);
    const int cn = (j0 + 1 < maxj ? j0 + 1 : maxj - 1) * maxi + i0;
    const int cs = (j0 > 0 ? j0 - 1 : 0) * maxi + i0;
    const R* mj = Met<R>::j(g);
    const R* mk = Met<R>::k(g);
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
    const int nl = (L - l0) < LB ? (L - l0) : LB;

    // carried values: [0]=c, [1]=e, [2]=w, [3]=n, [4]=s at level k (cur) and k+1 (up)
    R cur[LB][5], up[LB][5], fb[LB];
    R tcur[2][5], tup[2][5];
#pragma unroll
    for (int q = 0; q < LB; q++) fb[q] = (R)0;
    int32_t limps = 0;
    R dmaxv = (R)0;
    bool have_cur = false;

    for (int kz = k1c - 1; kz < maxk; kz++) {
        const int k = kz + 1;
        const size_t o = (size_t)kz * plane;
        if (!have_cur) {
#pragma unroll
            for (int q = 0; q < LB; q++) {
                if (q < nl) {
                    const R* s = src + (size_t)(l0 + q) * tstride + o;
                    cur[q][0] = s[col]; cur[q][1] = s[ce]; cur[q][2] = s[cw]; cur[q][3] = s[cn]; cur[q][4] = s[cs];
                }
            }
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const R* s = src + (size_t)t * tstride + o;
                tcur[t][0] = s[col]; tcur[t][1] = s[ce]; tcur[t][2] = s[cw]; tcur[t][3] = s[cn]; tcur[t][4] = s[cs];
            }
            have_cur = true;
        }
        const bool top = (k == maxk);
        if (!top) {
            const size_t ou = o + plane;
#pragma unroll
            for (int q = 0; q < LB; q++) {
                if (q < nl) {
                    const R* s = src + (size_t)(l0 + q) * tstride + ou;
                    up[q][0] = s[col]; up[q][1] = s[ce]; up[q][2] = s[cw]; up[q][3] = s[cn]; up[q][4] = s[cs];
                }
            }
#pragma unroll
            for (int t = 0; t < 2; t++) {
                const R* s = src + (size_t)t * tstride + ou;
                tup[t][0] = s[col]; tup[t][1] = s[ce]; tup[t][2] = s[cw]; tup[t][3] = s[cn]; tup[t][4] = s[cs];
            }
        }
        // face velocities and Peclet-weighted upwinding (2489-2523)
        const R ue_ = u[o + col], uw_ = u[o + cw];
        const R un_ = u[plane * maxk + o + col], us_ = u[plane * maxk + o + cs];
        const R ua_ = u[2 * plane * maxk + o + col];
        const R dza = mk[MK_DZA * KS + k], rdza = mk[MK_RDZA * KS + k];
        const R dt = mk[MK_DT * KS + k], rdz = mk[MK_RDZ * KS + k];
        R pec;
        pec = ue_ * dphi / diff1; const R upe = pec / ((R)2.0 + fabs(pec));
        pec = uw_ * dphi / diff1; const R upw = pec / ((R)2.0 + fabs(pec));
        pec = un_ * dsvn / diff1; const R upn = pec / ((R)2.0 + fabs(pec));
        pec = us_ * dsvs / diff1; const R ups = pec / ((R)2.0 + fabs(pec));
        pec = ua_ * dza / diffv; const R upa = pec / ((R)2.0 + fabs(pec));
        const bool oe = k >= max(k1c, k1e), ow = k >= max(k1w, k1c);
        const bool on = k >= max(k1c, k1n), os = k >= max(k1s, k1c);

        // isoneutral coefficients from T and S (2489-2501, 2561-2605)
        bool iso = false;
        R dzrho = 0, slim = 1, dxr[4] = {0, 0, 0, 0}, dyr[4] = {0, 0, 0, 0};
        if (!top) {
            const R tatw = (R)0.5 * (tcur[0][0] + tup[0][0]);
            const R tec = -e0 - e2 * tatw * 2 - e3 * tatw * tatw * 3;
            dzrho = (e1 * (tup[1][0] - tcur[1][0]) - tec * (tup[0][0] - tcur[0][0])) * rdza;
            if (M.diso && dzrho < (R)-1.0E-12) {
                iso = true;
                const R rdzrho = (R)1.0 / dzrho;
                R tv1 = 0;
                R dxT[4], dxS[4], dyT[4], dyS[4];
                // ina order: (knp,nnp) = (0,0), (0,1), (1,0), (1,1)
                dxT[0] = k >= k1w ? (tcur[0][0] - tcur[0][2]) * rc * rdphi : (R)0;
                dxS[0] = k >= k1w ? (tcur[1][0] - tcur[1][2]) * rc * rdphi : (R)0;
                dyT[0] = k >= k1s ? (tcur[0][0] - tcur[0][4]) * cvs * rdsvs : (R)0;
                dyS[0] = k >= k1s ? (tcur[1][0] - tcur[1][4]) * cvs * rdsvs : (R)0;
                dxT[1] = k >= k1e ? (tcur[0][1] - tcur[0][0]) * rc * rdphi : (R)0;
                dxS[1] = k >= k1e ? (tcur[1][1] - tcur[1][0]) * rc * rdphi : (R)0;
                dyT[1] = k >= k1n ? (tcur[0][3] - tcur[0][0]) * cvn * rdsvn : (R)0;
                dyS[1] = k >= k1n ? (tcur[1][3] - tcur[1][0]) * cvn * rdsvn : (R)0;
                dxT[2] = k + 1 >= k1w ? (tup[0][0] - tup[0][2]) * rc * rdphi : (R)0;
                dxS[2] = k + 1 >= k1w ? (tup[1][0] - tup[1][2]) * rc * rdphi : (R)0;
                dyT[2] = k + 1 >= k1s ? (tup[0][0] - tup[0][4]) * cvs * rdsvs : (R)0;
                dyS[2] = k + 1 >= k1s ? (tup[1][0] - tup[1][4]) * cvs * rdsvs : (R)0;
                dxT[3] = k + 1 >= k1e ? (tup[0][1] - tup[0][0]) * rc * rdphi : (R)0;
                dxS[3] = k + 1 >= k1e ? (tup[1][1] - tup[1][0]) * rc * rdphi : (R)0;
                dyT[3] = k + 1 >= k1n ? (tup[0][3] - tup[0][0]) * cvn * rdsvn : (R)0;
                dyS[3] = k + 1 >= k1n ? (tup[1][3] - tup[1][0]) * cvn * rdsvn : (R)0;
#pragma unroll
                for (int n = 0; n < 4; n++) {
                    dxr[n] = scc * dxS[n] - tec * dxT[n];
                    dyr[n] = scc * dyS[n] - tec * dyT[n];
                    tv1 = tv1 + dxr[n] * dxr[n] + dyr[n] * dyr[n];
                }
                tv1 = (R)0.25 * tv1 * rdzrho * rdzrho;
                const R ssmax = mk[MK_SSMAX * KS + k];
                if (tv1 > ssmax) {
                    slim = ssmax * ssmax / (tv1 * tv1);
                    limps++;
                } else {
                    slim = 1;
                }
                tv1 = tv1 * slim * diff1 * rdza;
                const R tvd = tv1 * dt * rdza;
                if (tvd > dmaxv) dmaxv = tvd;
            }
        }
        double newT = 0, newS = 0;
#pragma unroll
        for (int q = 0; q < LB; q++) {
            if (q >= nl) continue;
            const R Tc = cur[q][0], Te = cur[q][1], Tw = cur[q][2], Tn = cur[q][3], Ts = cur[q][4];
            const R fe = oe ? ue_ * rc * (((R)1.0 - upe) * Te + ((R)1.0 + upe) * Tc) * (R)0.5 - (Te - Tc) * rc2 * diff1
                            : (R)0;
            const R fw = ow ? uw_ * rc * (((R)1.0 - upw) * Tc + ((R)1.0 + upw) * Tw) * (R)0.5 - (Tc - Tw) * rc2 * diff1
                            : (R)0;
            const R fn = on ? cvn * un_ * (((R)1.0 - upn) * Tn + ((R)1.0 + upn) * Tc) * (R)0.5 - cv2n * (Tn - Tc) * diff1
                            : (R)0;
            const R fs = os ? cvs * us_ * (((R)1.0 - ups) * Tc + ((R)1.0 + ups) * Ts) * (R)0.5 - cv2s * (Tc - Ts) * diff1
                            : (R)0;
            R fa;
            if (top) {
                fa = M.sflux ? (R)M.sflux[(size_t)(l0 + q) * plane + col] : (R)0;
            } else {
                const R Tu = up[q][0];
                fa = ua_ * (((R)1.0 - upa) * Tu + ((R)1.0 + upa) * Tc) * (R)0.5;
                fa = fa - (Tu - Tc) * rdza * diffv;
                if (iso) {
                    const R dzts = (Tu - Tc) * rdza;
                    R dx[4], dy[4];
                    dx[0] = k >= k1w ? (Tc - Tw) * rc * rdphi : (R)0;
                    dy[0] = k >= k1s ? (Tc - Ts) * cvs * rdsvs : (R)0;
                    dx[1] = k >= k1e ? (Te - Tc) * rc * rdphi : (R)0;
                    dy[1] = k >= k1n ? (Tn - Tc) * cvn * rdsvn : (R)0;
                    dx[2] = k + 1 >= k1w ? (up[q][0] - up[q][2]) * rc * rdphi : (R)0;
                    dy[2] = k + 1 >= k1s ? (up[q][0] - up[q][4]) * cvs * rdsvs : (R)0;
                    dx[3] = k + 1 >= k1e ? (up[q][1] - up[q][0]) * rc * rdphi : (R)0;
                    dy[3] = k + 1 >= k1n ? (up[q][3] - up[q][0]) * cvn * rdsvn : (R)0;
                    R tv = 0;
#pragma unroll
                    for (int n = 0; n < 4; n++)
                        tv = tv + ((R)2 * dzrho * dx[n] - dxr[n] * dzts) * dxr[n] +
                             ((R)2 * dzrho * dy[n] - dyr[n] * dzts) * dyr[n];
                    tv = (R)0.25 * slim * diff1 * tv / (dzrho * dzrho);
                    fa = fa + tv;
                }
            }
            const R div = ((fe - fw) * rdphi + (fn - fs) * rds) + (fa - fb[q]) * rdz;
            const size_t idx = (size_t)(l0 + q) * tstride + o + col;
            double nv;
            if (sizeof(R) == 4 && lo_src) {
                // compensated accumulation: the rounding residual of every
                // update is carried to the next step in the int16 plane
                nv = StateIO<R>::ld(src, lo_src, idx) + (double)(-(dt * div));
                StateIO<R>::st(dst, lo_dst, idx, nv);
            } else {
                const R v = Tc - dt * div;
                dst[idx] = v;
                nv = (double)v;
            }
            fb[q] = fa;
            if (l0 + q == 0) newT = nv;
            if (l0 + q == 1) newS = nv;
        }
        if (l0 == 0) M.rho[o + col] = eos<double>(g.ec[0], g.ec[1], g.ec[2], g.ec[3], newT, newS);
        // carry level k+1 down
        if (!top) {
#pragma unroll
            for (int q = 0; q < LB; q++)
#pragma unroll
                for (int n = 0; n < 5; n++) cur[q][n] = up[q][n];
#pragma unroll
            for (int t = 0; t < 2; t++)
#pragma unroll
                for (int n = 0; n < 5; n++) tcur[t][n] = tup[t][n];
        }
    }
    if (l0 == 0 && M.diso) {
        if (limps) atomicAdd(&M.diag[0], limps);
        if (dmaxv > (R)0) atomic_max_pos(M.dmax, (double)dmaxv);
    }
}

// co, goldstein.f90:2657-2777 (iconv = 0, ieos = 0): one thread per wet
// column of a member, FP64 arithmetic on the new state in place.
template <typename R>
__global__ __launch_bounds__(64) void k_co(const MemberDev* __restrict__ mems, const int32_t* __restrict__ list) {
    const MemberDev& M = mems[list[blockIdx.y]];
    const GridDev& g = M.g;
    const int w = blockIdx.x * 64 + threadIdx.x;
    if (w >= g.nwet) return;
    const int col = g.wetcols[w];
    const int maxi = g.maxi, maxk = g.maxk, L = M.L;
    const int I = col % maxi + 1, J = col / maxi + 1;
    const int k1 = k1at(g, I, J);
    const int par = *M.parity;
    void* st = M.st[par ^ 1];
    int16_t* lo = M.lo[par ^ 1];
    const size_t plane = (size_t)g.ncol, tstride = (size_t)maxk * plane;
    double* rho = M.rho;
    const double* dz = g.mk + MK_DZ * (maxk + 2);
    const double e0 = g.ec[0], e1 = g.ec[1], e2 = g.ec[2], e3 = g.ec[3];
    int kk[MAXK_DEV + 2];
    double dzm[MAXK_DEV + 2];
    double r[MAXK_DEV + 2];
    for (int m = 0; m < MAXK_DEV + 2; m++) kk[m] = 0;
    for (int m = k1; m <= maxk; m++) r[m] = rho[(size_t)(m - 1) * plane + col];
    kk[k1 - 1] = 0;
    for (int m = k1; m <= maxk; m++) {
        kk[m] = m;
        dzm[m] = dz[m];
    }
    int m = maxk, lastmix = 0;
    bool mixed = false;
    // every pass either moves m or merges two segments, so the reference loop
    // ends within O(maxk^2) passes; the cap only guards non-finite densities
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
            int n = m - 1;
            while (kk[n - 1] > 0 && r[kk[n]] >= r[kk[n - 1]]) n = n - 1;
            const int km = kk[m];
            for (int l = 0; l < L; l++) {
                const size_t base = (size_t)l * tstride + col;
                double s = StateIO<R>::ld(st, lo, base + (size_t)(km - 1) * plane) * dzm[km];
                for (int ni = 1; ni <= m - n; ni++)
                    s = s + StateIO<R>::ld(st, lo, base + (size_t)(kk[m - ni] - 1) * plane) * dzm[kk[m - ni]];
                double dsum = dzm[km];
                for (int ni = 1; ni <= m - n; ni++) dsum = dsum + dzm[kk[m - ni]];
                StateIO<R>::st(st, lo, base + (size_t)(km - 1) * plane, s / dsum);
            }
            for (int ni = 1; ni <= m - n; ni++) dzm[km] = dzm[km] + dzm[kk[m - ni]];
            const double t = StateIO<R>::ld(st, lo, (size_t)(km - 1) * plane + col);
            const double sl = StateIO<R>::ld(st, lo, tstride + (size_t)(km - 1) * plane + col);
            r[km] = eos<double>(e0, e1, e2, e3, t, sl);
            int ni = m - 1;
            while (kk[ni + 1] > 0) {
                kk[ni] = kk[ni - m + n];
                ni = ni - 1;
            }
        }
    }
    if (guard <= 0) atomicAdd(&M.diag[3], 1);
    m = maxk - 1;
    double cost = 0.0;
    for (int n = maxk - 1; n >= k1; n--) {
        if (n > kk[m]) {
            for (int l = 0; l < L; l++) {
                const size_t base = (size_t)l * tstride + col;
                StateIO<R>::st(st, lo, base + (size_t)(n - 1) * plane,
                               StateIO<R>::ld(st, lo, base + (size_t)(kk[m + 1] - 1) * plane));
            }
            const double t = StateIO<R>::ld(st, lo, (size_t)(n - 1) * plane + col);
            const double sl = StateIO<R>::ld(st, lo, tstride + (size_t)(n - 1) * plane + col);
            r[n] = eos<double>(e0, e1, e2, e3, t, sl);
            cost += 1.0;
        } else {
            m = m - 1;
        }
    }
    if (mixed || cost > 0.0) {
        for (int q = k1; q <= maxk; q++) rho[(size_t)(q - 1) * plane + col] = r[q];
        if (cost > 0.0) M.cost[col] += cost;
    }
}

#ifdef CGO_DEFINE_FLIP
// end of step: flip the ping-pong parity, count the step, roll the mean salinity
__global__ void k_flip(const MemberDev* __restrict__ mems, const int32_t* __restrict__ list, int n) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const MemberDev& M = mems[list[t]];
    *M.parity ^= 1;
    *M.steps += 1;
    if (M.bgc) M.scal[SC_SMEAN] = M.scal[SC_SNEW];
}
#endif

}  // namespace cgo
