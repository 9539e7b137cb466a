// k_mom.hip -- GOLDSTEIN's velocity solve on the device: the barotropic
// streamfunction with islands and the baroclinic velocities of every member
// with a wind stress (goldstein.f90:198-233: wind 3685-3714, jbar 3217-3315,
// ubarsolv 3500-3565, the island correction with island() of
// goldstein_lib.f90:186-241 and matmult 3470-3492, velc 3568-3679).
//
// Four launches per step over the members with a solve:
//   k_mom_cols_a  per column, many blocks per member: jbar's bottom pressure
//                 and its integral, and velc's thermal-wind (and top-layer
//                 Ekman) shear integrated up the column with its depth sum
//   k_mom_solve   one block per member: the psi forcing (wind + JEBAR), the
//                 banded substitutions, psi and the barotropic velocities,
//                 the path integrals round the islands and psibc
//   k_mom_cols_b  velc's barotropic part and relaxation at every open face
//   k_mom_cols_c  w from continuity and, for an FP32 member, the seven flux
//                 coefficient planes formed from the FP64 velocities; the
//                 island correction of the stored ub and psi
// The substitutions are chains of dependent operations (elimination step i
// needs row i final; back substitution subtracts a row's terms in the
// reference's order): one wave runs them while the block's other three waves
// stream the factor rows from memory into a ring in LDS a chunk ahead.
// Every operation rounds as written (-ffp-contract=off) in the reference's
// order, so an FP64 member's velocities are the reference arithmetic's.
#include "cgocn_dev.h"

namespace cgo {

constexpr int MOM_NT = 256;    // threads of every block
constexpr int MOM_CH = 32;     // factor rows per chunk of the LDS ring
constexpr int MOM_RING = 3;    // chunks the ring holds: the one in use, the next, the one loading
constexpr int MOM_PRE = 8;     // elimination steps whose multipliers are read from LDS ahead

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ int uni(int v) { return __builtin_amdgcn_readfirstlane(v); }

// the arrays of one member's solve, addressed with the reference's indices
struct MomView {
    const MemberDev& M;
    const MomGridDev& G;
    int mi, mj, mk, n, N, ncol, ncell;
    MomMemberLayout ML;
    double* W;
    double* U;
    __device__ explicit MomView(const MemberDev& m)
        : M(m), G(*m.mg), mi(uni(G.maxi)), mj(uni(G.maxj)), mk(uni(G.maxk)), n(uni(G.n)), N(uni(G.N)),
          ncol(mi * mj), ncell(mi * mj * mk), ML(mom_member_layout(mi, mj, mk, m.prec != 0)), W(m.mw), U(m.u64) {}
    __device__ int K1(int i, int j) const { return k1at(M.g, i, j); }
    __device__ double jt(int t, int j) const { return G.jt[t * (mj + 2) + j]; }
    __device__ double rh(int l, int i, int j) const { return G.rh[(l - 1) + 3 * (i + (mi + 2) * j)]; }
    __device__ double drag(int l, int i, int j) const { return G.drag[(l - 1) + 2 * (i + (mi + 2) * j)]; }
    __device__ int ku(int l, int i, int j) const { return G.ku[(l - 1) + 2 * ((i - 1) + mi * (j - 1))]; }
    __device__ int mkk(int i, int j) const { return G.mk[i + (mi + 2) * j]; }
    __device__ int col(int i, int j) const { return (i - 1) + mi * (j - 1); }
    // rho at a wet cell, i = 0..maxi+1 periodic
    __device__ double rho(int i, int j, int k) const {
        const int ii = i == 0 ? mi : (i == mi + 1 ? 1 : i);
        return M.rho[(size_t)(k - 1) * ncol + col(ii, j)];
    }
    __device__ double& bp(int i, int j, int k) const { return W[ML.bp + i + (mi + 2) * (j + (mj + 2) * k)]; }
    __device__ double& sbp(int i, int j) const { return W[ML.sbp + i + (mi + 2) * j]; }
    __device__ double& ub(int l, int i, int j) const { return W[ML.ub + (l - 1) + 2 * (i + (mi + 2) * j)]; }
    __device__ double& psi(int i, int j) const { return W[ML.psi + i + (mi + 1) * j]; }
    __device__ double tau(int l, int i, int j) const { return W[ML.tau + (l - 1) + 2 * col(i, j)]; }
    __device__ double dztau(int l, int i, int j) const { return W[ML.dztau + (l - 1) + 2 * col(i, j)]; }
    __device__ double dztav(int l, int i, int j) const { return W[ML.dztav + (l - 1) + 2 * col(i, j)]; }
    // u, v, w at interior points (i = 1..maxi, j = 1..maxj)
    __device__ double& uu(int l, int i, int j, int k) const {
        return U[(size_t)(l - 1) * ncell + (size_t)(k - 1) * ncol + col(i, j)];
    }
    __device__ double& u1(int l, int i, int j, int k) const {
        return W[ML.u1 + (size_t)(l - 1) * ncell + (size_t)(k - 1) * ncol + col(i, j)];
    }
    __device__ double& csum(int l, int i, int j) const { return W[ML.csum + (size_t)(l - 1) * ncol + col(i, j)]; }
    __device__ double psibc(int isl) const { return W[ML.psibc + isl - 1]; }
};

// ------------------------------------------------------------ per column, before the solve
__global__ __launch_bounds__(MOM_NT) void k_mom_cols_a(const MemberDev* __restrict__ mems,
                                                       const int32_t* __restrict__ list) {
    const MomView V(mems[list[blockIdx.y]]);
    const int c = blockIdx.x * MOM_NT + threadIdx.x;
    if (c >= V.ncol) return;
    const int mi = V.mi, mk = V.mk, i = c % mi + 1, j = c / mi + 1, kb = V.K1(i, j);
    if (kb > mk) return;
    const MomGridDev& G = V.G;
    // jbar (3225-3255): bottom pressure from the bottom cell up (zero there,
    // never written), the periodic copy, and its integral above mk, summed as
    // the levels are formed (mk(i,j) >= k1(i,j), so every term is one of them)
    const int mkij = V.mkk(i, j);
    double b = V.bp(i, j, kb);
    double sb = 0;
    for (int k = kb + 1; k <= mk; k++) {
        b = b - (V.rho(i, j, k) + V.rho(i, j, k - 1)) * G.dza[k - 1] * 0.5;
        V.bp(i, j, k) = b;
        if (i == 1 && kb < mk) V.bp(mi + 1, j, k) = b;
        if (mkij > 0 && k >= mkij + 1) sb = sb + b * G.dz[k];
    }
    if (mkij > 0) V.sbp(i, j) = sb;
    if (i == 1) V.sbp(mi + 1, j) = V.sbp(1, j);
    // velc (3574-3640): the shear integrated up from zero at the bottom
    double sum0 = 0, sum1 = 0, dp0 = 0, dp1 = 0, up0 = 0, up1 = 0;
    const double sj = V.jt(JT_S, j), svj = V.jt(JT_SV, j), cj = V.jt(JT_C, j), cvj = V.jt(JT_CV, j);
    const double rdphi = G.rdphi, rtvij = G.rtv[V.col(i, j)], rtv3ij = G.rtv3[V.col(i, j)];
    const double d1 = V.drag(1, i, j), d2 = V.drag(2, i, j);
    const double jrc = V.jt(JT_RC, j), jrds2 = V.jt(JT_RDS2, j), jrdsvm = V.jt(JT_RDSV, j - 1), jrdsv = V.jt(JT_RDSV, j),
                 jrcv = V.jt(JT_RCV, j);
    // the neighbours' bottom levels, once
    const int kE = V.K1(i + 1, j), kN = V.K1(i, j + 1), kS = V.K1(i, j - 1), kNE = V.K1(i + 1, j + 1),
              kSE = V.K1(i + 1, j - 1), kW = V.K1(i - 1, j), kNW = V.K1(i - 1, j + 1);
    // the eight densities a level reads: (i,j) (i+1,j) (i+1,j+1) (i+1,j-1)
    // (i,j+1) (i,j-1) (i-1,j+1) (i-1,j), the next level's fetched while one
    // level is formed; rows are clamped to the grid for the fetch, and a
    // value is used only where its row is wet, so the clamp never shows
    const int jp = min(j + 1, V.mj), jm = max(j - 1, 1);
    auto fetch = [&](int k, double* r) {
        r[0] = V.rho(i, j, k);
        r[1] = V.rho(i + 1, j, k);
        r[2] = V.rho(i + 1, jp, k);
        r[3] = V.rho(i + 1, jm, k);
        r[4] = V.rho(i, jp, k);
        r[5] = V.rho(i, jm, k);
        r[6] = V.rho(i - 1, jp, k);
        r[7] = V.rho(i - 1, j, k);
    };
    double cur[8], nxt[8];
    fetch(kb, cur);
    for (int k = kb; k <= mk; k++) {
        if (k < mk) fetch(k + 1, nxt);
        const double r_c = cur[0], r_e = cur[1], r_ne = cur[2], r_se = cur[3], r_n = cur[4], r_s = cur[5],
                     r_nw = cur[6], r_w = cur[7];
        double tv1, tv2, tv4, tv5;
        if (kE > k) {
            tv1 = 0;
            tv2 = 0;
        } else {
            tv2 = -(r_e - r_c) * rdphi * jrc;
            if (max(max(kS, kN), max(kSE, kNE)) <= k)
                tv1 = -cj * (r_ne - r_se + r_n - r_s) * jrds2 * 0.25;
            else if (max(kS, kSE) <= k)
                tv1 = -cj * (r_e - r_se + r_c - r_s) * jrdsvm * 0.5;
            else if (max(kN, kNE) <= k)
                tv1 = -cj * (r_ne - r_e + r_n - r_c) * jrdsv * 0.5;
            else
                tv1 = 0;
        }
        if (kN > k) {
            tv4 = 0;
            tv5 = 0;
        } else {
            tv4 = -cvj * (r_n - r_c) * jrdsv;
            if (max(max(kW, kNW), max(kE, kNE)) <= k)
                tv5 = -(r_ne - r_nw + r_e - r_w) * rdphi * 0.25 * jrcv;
            else if (max(kW, kNW) <= k)
                tv5 = -(r_n - r_nw + r_c - r_w) * rdphi * 0.5 * jrcv;
            else if (max(kE, kNE) <= k)
                tv5 = -(r_ne - r_n + r_e - r_c) * rdphi * 0.5 * jrcv;
            else
                tv5 = 0;
        }
        if (k == mk) {
            if (kE <= k) {
                tv1 = tv1 - V.dztau(2, i, j);
                tv2 = tv2 - V.dztau(1, i, j);
            }
            if (kN <= k) {
                tv4 = tv4 - V.dztav(2, i, j);
                tv5 = tv5 - V.dztav(1, i, j);
            }
        }
        const double dz0 = -(sj * tv1 + d1 * tv2) * rtvij;
        const double dz1 = -(d2 * tv4 - svj * tv5) * rtv3ij;
        double x0, x1;
        if (k == kb) {
            x0 = 0;
            x1 = 0;
        } else {
            x0 = up0 + G.dza[k - 1] * (dz0 + dp0) * 0.5;
            sum0 = sum0 + G.dz[k] * x0;
            x1 = up1 + G.dza[k - 1] * (dz1 + dp1) * 0.5;
            sum1 = sum1 + G.dz[k] * x1;
        }
        V.uu(1, i, j, k) = x0;
        V.uu(2, i, j, k) = x1;
        up0 = x0;
        up1 = x1;
        dp0 = dz0;
        dp1 = dz1;
#pragma unroll
        for (int q = 0; q < 8; q++) cur[q] = nxt[q];
    }
    V.csum(1, i, j) = sum0;
    V.csum(2, i, j) = sum1;
}

// ------------------------------------------------------------ the band solve
// ubarsolv's substitutions (goldstein.f90:3508-3522) on the psi forcing g in
// LDS (row i at g[i-1]), which ends as the solution.  Called by the whole
// block: wave 0 runs the chains, waves 1-3 fill the ring with factor rows
// one chunk ahead, and the block synchronises once per chunk.
//
// Forward: elimination step i subtracts ratm(i+q, q) x_i from rows i+1..i+n+1.
// Lane (j-1) mod 64 keeps row j in a register from the step it first changes
// until it becomes the pivot (a row is live for at most n+2 < 64 steps, so
// the 64 lanes hold every live row); step i reads the pivot from its lane and
// every lane holding a row in the band updates it.  Ring row r of a chunk
// holds the multipliers of one step.
//
// Back substitution: row i subtracts gap(i, n+2+k) x_{i+k} for k = 1, 2, ..
// in that order, then divides by the diagonal.  A term with a zero
// coefficient subtracts a signed zero and leaves the value as it is, so only
// the others are taken (bit k-1 of the row's mask): the k = 1 term from the
// row just solved, and the products for k >= 2, which need only rows solved
// earlier, formed one row ahead by the lanes and packed in order into a
// zero-padded LDS buffer that the chain reads.  Dividing by a diagonal of
// exactly 1 (land points) is the identity and is skipped.  Ring row r of a
// chunk holds a row of the eliminated band (n+2 values) and its mask.
__device__ __forceinline__ void band_solve(double* g, double* pbuf, double* ring, const MomGridDev& G, int tid) {
    const int n = uni(G.n), N = uni(G.N), W1 = n + 1, W2 = n + 2, RS = n + 3;
    const int lane = tid & 63;
    const double* __restrict__ fwd = G.fwd;
    const double* __restrict__ bwd = G.bwd;
    const uint64_t* __restrict__ bmask = G.bmask;
    constexpr int PT = 12;  // ring values a producer thread moves per chunk (>= 32 x 64 / 192)
    auto slot = [&](int c) { return ring + (size_t)(c % MOM_RING) * MOM_CH * RS; };
    // producers: the multipliers of steps [c CH + 1, c CH + CH]
    auto produce_f = [&](int c) {
        double* dst = slot(c);
        const int r0 = c * MOM_CH, nr = min(MOM_CH, N - 1 - r0), tot = nr * W1;
        double t[PT];
#pragma unroll
        for (int e = 0; e < PT; e++) {
            const int x = (tid - 64) + e * (MOM_NT - 64);
            t[e] = x < tot ? fwd[(size_t)r0 * W1 + x] : 0.0;
        }
#pragma unroll
        for (int e = 0; e < PT; e++) {
            const int x = (tid - 64) + e * (MOM_NT - 64);
            if (x < tot) dst[(x / W1) * RS + x % W1] = t[e];
        }
    };
    // producers: rows i = N-1 - c CH - t of the eliminated band and their masks
    auto produce_b = [&](int c) {
        double* dst = slot(c);
        const int ihi = N - 1 - c * MOM_CH, nr = min(MOM_CH, ihi), tot = nr * (W2 + 1);
        double t[PT + 1];
#pragma unroll
        for (int e = 0; e < PT + 1; e++) {
            const int x = (tid - 64) + e * (MOM_NT - 64);
            const int r = x / (W2 + 1), k = x % (W2 + 1), i = ihi - r;
            double val = 0.0;
            if (x < tot) {
                if (k < W2)
                    val = bwd[(size_t)(i - 1) * W2 + k];
                else
                    val = __longlong_as_double((long long)bmask[i - 1]);
            }
            t[e] = val;
        }
#pragma unroll
        for (int e = 0; e < PT + 1; e++) {
            const int x = (tid - 64) + e * (MOM_NT - 64);
            if (x < tot) dst[(x / (W2 + 1)) * RS + x % (W2 + 1)] = t[e];
        }
    };

    // ---- forward elimination
    const int nchf = uni((N - 1 + MOM_CH - 1) / MOM_CH);
    if (tid >= 64) {
        produce_f(0);
        if (nchf > 1) produce_f(1);
    }
    double v = 0.0, vn = 0.0;
    if (tid < 64) v = lane < N ? g[lane] : 0.0;
    __syncthreads();
    for (int c = 0; c < nchf; c++) {
        c = uni(c);
        if (tid < 64) {
            const double* rr = slot(c);
            const int sb = (c & ~1) * MOM_CH;             // the 64-step block's first step, minus one
            const int pb = (c & 1) * MOM_CH;              // the pivot lane of the chunk's first step
            if ((c & 1) == 0) vn = g[min(sb + lane + 64, N - 1)];   // row sb+L+65's forcing, untouched so far
            double r[MOM_PRE];
#pragma unroll
            for (int d = 0; d < MOM_PRE; d++) {
                const int q = (lane - (pb + d)) & 63;
                r[d] = rr[d * RS + (q >= 1 && q <= n + 1 ? q - 1 : 0)];
            }
#pragma unroll
            for (int d = 0; d < MOM_CH; d++) {
                const int p = pb + d, s = c * MOM_CH + d + 1;
                const double rd = r[d % MOM_PRE];
                if (d + MOM_PRE < MOM_CH) {
                    const int q2 = (lane - (p + MOM_PRE)) & 63;
                    r[d % MOM_PRE] = rr[(d + MOM_PRE) * RS + (q2 >= 1 && q2 <= n + 1 ? q2 - 1 : 0)];
                }
                if (s <= N - 1) {
                    const double x = readlane_d(v, p);
                    const int q = (lane - p) & 63;
                    const bool act = q >= 1 && q <= n + 1 && s + q <= N;
                    const double upd = v - rd * x;
                    v = act ? upd : v;
                    if (lane == 0) g[s - 1] = x;
                    if (lane == p) v = sb + lane + 65 <= N ? vn : 0.0;
                }
            }
        } else if (c + 2 < nchf) {
            produce_f(c + 2);
        }
        __syncthreads();
    }
    // ---- back substitution
    const int nchb = uni((N - 1 + MOM_CH - 1) / MOM_CH);   // rows N-1 .. 1
    if (tid >= 64) {
        produce_b(0);
        if (nchb > 1) produce_b(1);
    }
    double xp = 0.0;
    if (tid < 64) {
        // row N is final after the last step: it sits in lane (N-1) mod 64
        xp = readlane_d(v, uni((N - 1) & 63));
        const double dN = bwd[(size_t)(N - 1) * W2];
        if (dN != 1.0) xp = xp / dN;
        if (lane == 0) g[N - 1] = xp;
        pbuf[lane] = 0.0;
        pbuf[64 + lane] = 0.0;
    }
    __syncthreads();
    const int ql = min(lane + 1, n + 1);
    // ring row of band row i
    auto brow = [&](int i) { return slot((N - 1 - i) / MOM_CH) + (size_t)((N - 1 - i) % MOM_CH) * RS; };
    auto maskof = [&](const double* r) { return (uint64_t)__double_as_longlong(r[W2]); };
    // carried from one row to the next: row i's mask, k = 1 coefficient,
    // diagonal and forward value; row i-1's mask and this lane's coefficient
    uint64_t m_i = 0, m_p = 0;
    double a1_i = 0.0, d_i = 1.0, g_i = 0.0, c_p = 0.0;
    if (tid < 64) {
        const double* r0 = brow(N - 1);
        m_i = maskof(r0);
        a1_i = r0[1];
        d_i = r0[0];
        g_i = g[N - 2];
        if (N - 2 >= 1) {
            const double* r1 = brow(N - 2);
            m_p = maskof(r1);
            c_p = r1[ql];
        }
    }
    for (int c = 0; c < nchb; c++) {
        c = uni(c);
        if (tid < 64) {
            const int ihi = N - 1 - c * MOM_CH, ilo = max(1, ihi - MOM_CH + 1);
            for (int i = ihi; i >= ilo; i--) {
                i = uni(i);
                const double* pc = pbuf + (i & 1) * 64;
                const int pad = uni((__popcll(m_i & ~(uint64_t)1) + 7) & ~7);
                double pa[8], pn[8];
#pragma unroll
                for (int e = 0; e < 8; e++) pa[e] = pc[e];
                // the next rows' data, read now and used in the next iteration
                uint64_t m_n = 0, m_pn = 0;
                double a1_n = 0.0, d_n = 1.0, g_n = 0.0, c_pn = 0.0;
                if (i >= 2) {
                    const double* rn = brow(i - 1);
                    m_n = maskof(rn);
                    a1_n = rn[1];
                    d_n = rn[0];
                    g_n = g[i - 2];
                }
                if (i >= 3) {
                    const double* rnn = brow(i - 2);
                    m_pn = maskof(rnn);
                    c_pn = rnn[ql];
                }
                const double xk = g[min(i + lane - 1, N - 1)];   // x_{i-1+q}, q = lane + 1 (rows i+1.. for q >= 2)
                // the chain of row i
                double t = g_i;
                if (m_i & 1) t = t - a1_i * xp;
                for (int q = 0; q < pad; q += 8) {
                    if (q + 8 < pad) {
#pragma unroll
                        for (int e = 0; e < 8; e++) pn[e] = pc[q + 8 + e];
                    }
#pragma unroll
                    for (int e = 0; e < 8; e++) t = t - pa[e];
#pragma unroll
                    for (int e = 0; e < 8; e++) pa[e] = pn[e];
                }
                if (d_i != 1.0) t = t / d_i;
                if (lane == 0) g[i - 1] = t;
                xp = t;
                // products of row i-1 for k >= 2 into its buffer
                if (i >= 2) {
                    const uint64_t m = m_p & ~(uint64_t)1;
                    const bool set = (m >> lane) & 1;
                    const int rank = __popcll(m & (((uint64_t)1 << lane) - 1));
                    const int cnt = __popcll(m), ppad = (cnt + 7) & ~7;
                    double* pb = pbuf + ((i - 1) & 1) * 64;
                    if (lane >= cnt && lane < ppad) pb[lane] = 0.0;   // the padding (rank < cnt: no overlap)
                    if (set) pb[rank] = c_p * xk;
                }
                m_i = m_n;
                a1_i = a1_n;
                d_i = d_n;
                g_i = g_n;
                m_p = m_pn;
                c_p = c_pn;
            }
        } else if (c + 2 < nchb) {
            produce_b(c + 2);
        }
        __syncthreads();
    }
}

// ------------------------------------------------------------ one block per member
// mode 0: the whole solve; the FP32 form's dense solve splits it round
// k_mom_gemm: mode 1 forms the psi forcing and stages it, mode 2 takes psi
// from the staging and goes on from there
__global__ __launch_bounds__(MOM_NT) void k_mom_solve(const MemberDev* __restrict__ mems,
                                                      const int32_t* __restrict__ list, int mode) {
    extern __shared__ double lds[];
    const MomView V(mems[list[blockIdx.x]]);
    const MomGridDev& G = V.G;
    const int mi = V.mi, mj = V.mj, n = V.n, N = V.N, isles = G.isles, tid = threadIdx.x;
    double* gbs = lds;                          // the psi forcing, then psi (N)
    double* erhs = lds + N;                     // island integrals, then psibc (isles)
    double* pbuf = erhs + MOM_MAXISL;           // back substitution's staged products (2 x 64)
    double* ring = pbuf + 128;                  // factor rows (MOM_RING x MOM_CH x (n+3)), then island terms
    const double rdphi = G.rdphi, dphi = G.dphi;

    // ---- the psi forcing: the wind part (wind) plus the torque of the bottom
    // pressure where the four columns round a psi point differ in depth
    // (jbar 3256-3313)
    for (int l = tid + 1; l <= N && mode != 2; l += MOM_NT) {
        const int i = (l - 1) % n + 1, j = (l - 1) / n;
        double g = V.W[V.ML.gbw + l - 1];
        if (j >= 1 && j <= mj - 1 && G.getj[V.col(i, j)]) {
            const int ip1 = i % mi + 1;
            auto span = [&](int k0, int k1_, int ii, int jj) {
                double t = 0;
                for (int k = k0; k <= k1_; k++) t = t + V.bp(ii, jj, k) * G.dz[k];
                return t;
            };
            double tv1 = span(V.ku(2, ip1, j), V.mkk(ip1, j + 1), ip1, j + 1);
            double tv2 = span(V.ku(2, ip1, j), V.mkk(ip1, j), ip1, j);
            double tv3 = span(V.ku(2, i, j), V.mkk(i, j + 1), i, j + 1);
            double tv4 = span(V.ku(2, i, j), V.mkk(i, j), i, j);
            g = g + ((tv3 + V.sbp(i, j + 1) - tv4 - V.sbp(i, j)) * V.rh(2, i, j) -
                     (tv1 + V.sbp(ip1, j + 1) - tv2 - V.sbp(ip1, j)) * V.rh(2, ip1, j)) *
                        rdphi * V.jt(JT_RDSV, j);
            tv1 = span(V.ku(1, i, j + 1), V.mkk(ip1, j + 1), ip1, j + 1);
            tv2 = span(V.ku(1, i, j), V.mkk(ip1, j), ip1, j);
            tv3 = span(V.ku(1, i, j + 1), V.mkk(i, j + 1), i, j + 1);
            tv4 = span(V.ku(1, i, j), V.mkk(i, j), i, j);
            g = g + ((tv1 + V.sbp(ip1, j + 1) - tv3 - V.sbp(i, j + 1)) * V.rh(1, i, j + 1) -
                     (tv2 + V.sbp(ip1, j) - tv4 - V.sbp(i, j)) * V.rh(1, i, j)) *
                        rdphi * V.jt(JT_RDSV, j);
        }
        gbs[l - 1] = g;
        if (mode == 1) V.W[V.ML.gst + l - 1] = g;
    }
    if (mode == 1) return;
    __syncthreads();

    if (mode == 0) {
        band_solve(gbs, pbuf, ring, G, tid);
    } else {
        for (int l = tid; l < N; l += MOM_NT) {
            double p = V.W[V.ML.gst + N + l];
            for (int z = 1; z < MOM_SPLITK; z++) p = p + V.W[V.ML.gst + (size_t)N * (1 + z) + l];
            gbs[l] = p;
        }
        __syncthreads();
    }

    // ---- psi and the barotropic velocities from the solution (3524-3565),
    // each boundary copy written by the thread that forms its source
    for (int l = tid + 1; l <= N; l += MOM_NT) {
        const int i = (l - 1) % n + 1, j = (l - 1) / n;
        const double p = gbs[l - 1];
        V.psi(i, j) = p;
        if (i == mi) V.psi(0, j) = p;
        if (j >= 1) {
            const double x = -V.rh(1, i, j) * V.jt(JT_C, j) * (p - gbs[l - 1 - n]) * V.jt(JT_RDS, j);
            V.ub(1, i, j) = x;
            if (i == mi) V.ub(1, 0, j) = x;
            if (i == 1) V.ub(1, mi + 1, j) = x;
        }
        double y;
        if (j >= 1 && j <= mj - 1) {
            const double pw = gbs[(i == 1 ? mi : i - 1) + j * n - 1];
            y = V.rh(2, i, j) * (p - pw) * V.jt(JT_RCV, j) * rdphi;
        } else {
            y = 0.0;
        }
        V.ub(2, i, j) = y;
        if (i == 1) V.ub(2, mi + 1, j) = y;
        if (i == mi) V.ub(2, 0, j) = y;
    }
    __syncthreads();
    if (isles == 0) return;

    // ---- the path integral round every island of the momentum balance with
    // wind and bottom-pressure torque (island, indj = 1): each point's two
    // terms formed in parallel into LDS, then added in the path's order
    int base = 0;
    for (int isl = 1; isl <= isles; isl++) {
        const int np = G.npi[isl - 1];
        for (int p = tid + 1; p <= np; p += MOM_NT) {
            const int q = (p - 1) + G.mpi * (isl - 1);
            const int lpi = G.lpisl[q], ipi = G.ipisl[q], jpi = G.jpisl[q];
            const int al = abs(lpi);
            double cor;
            if (al == 1)
                cor = -V.jt(JT_S, jpi) * 0.25 *
                      (V.ub(2, ipi, jpi) + V.ub(2, ipi + 1, jpi) + V.ub(2, ipi, jpi - 1) + V.ub(2, ipi + 1, jpi - 1));
            else
                cor = V.jt(JT_SV, jpi) * 0.25 *
                      (V.ub(1, ipi - 1, jpi) + V.ub(1, ipi, jpi) + V.ub(1, ipi - 1, jpi + 1) + V.ub(1, ipi, jpi + 1));
            const double sgn = lpi >= 0 ? 1.0 : -1.0;
            const double rhv = V.rh(al, ipi, jpi), tauv = V.tau(al, ipi, jpi);
            ring[base + 2 * (p - 1)] = sgn * (V.drag(al, ipi, jpi) * V.ub(al, ipi, jpi) + cor - 1.0 * tauv * rhv) *
                                       (V.jt(JT_C, jpi) * dphi * (2.0 - al) + V.jt(JT_RCV, jpi) * V.jt(JT_DSV, jpi) * (al - 1.0));
            double t = 0.0, b2;
            if (al == 1) {
                for (int k = V.ku(1, ipi, jpi); k <= V.mkk(ipi + 1, jpi); k++) t = t + V.bp(ipi + 1, jpi, k) * G.dz[k];
                for (int k = V.ku(1, ipi, jpi); k <= V.mkk(ipi, jpi); k++) t = t - V.bp(ipi, jpi, k) * G.dz[k];
                b2 = (V.sbp(ipi + 1, jpi) - V.sbp(ipi, jpi) + t) * sgn * V.rh(1, ipi, jpi);
            } else {
                for (int k = V.ku(2, ipi, jpi); k <= V.mkk(ipi, jpi + 1); k++) t = t + V.bp(ipi, jpi + 1, k) * G.dz[k];
                for (int k = V.ku(2, ipi, jpi); k <= V.mkk(ipi, jpi); k++) t = t - V.bp(ipi, jpi, k) * G.dz[k];
                b2 = (V.sbp(ipi, jpi + 1) - V.sbp(ipi, jpi) + t) * sgn * V.rh(2, ipi, jpi);
            }
            ring[base + 2 * (p - 1) + 1] = b2;
        }
        base += 2 * np;
    }
    __syncthreads();
    if (tid == 0) {
        int b = 0;
        for (int isl = 1; isl <= isles; isl++) {
            double e = 0.0;
            for (int x = 0; x < 2 * G.npi[isl - 1]; x++) e = e + ring[b + x];
            erhs[isl - 1] = e;
            b += 2 * G.npi[isl - 1];
        }
        // the island values that cancel them (209-218); matmult with the
        // eliminated matrix of matinv_gold
        auto A = [&](int a, int b2) { return G.erisl[(a - 1) + isles * (b2 - 1)]; };
        if (isles > 1) {
            for (int i = 1; i <= isles - 1; i++)
                for (int j = i + 1; j <= isles; j++) erhs[j - 1] = A(i, i) * erhs[j - 1] - A(j, i) * erhs[i - 1];
            erhs[isles - 1] = erhs[isles - 1] / A(isles, isles);
            for (int i = isles - 1; i >= 1; i--) {
                for (int j = i + 1; j <= isles; j++) erhs[i - 1] = erhs[i - 1] - A(i, j) * erhs[j - 1];
                erhs[i - 1] = erhs[i - 1] / A(i, i);
            }
            for (int isl = 1; isl <= isles; isl++) V.W[V.ML.psibc + isl - 1] = -erhs[isl - 1];
        } else {
            V.W[V.ML.psibc] = -erhs[0] / A(1, 1);
        }
    }
}

// ------------------------------------------------------------ per column, after the solve
// velc (3641-3657): the barotropic velocity, with the island fields times
// their psi values added (219-227), in place of the shear's depth mean at
// every open face, relaxed towards the last step's
__global__ __launch_bounds__(MOM_NT) void k_mom_cols_b(const MemberDev* __restrict__ mems,
                                                       const int32_t* __restrict__ list) {
    const MomView V(mems[list[blockIdx.y]]);
    const MomGridDev& G = V.G;
    const int c = blockIdx.x * MOM_NT + threadIdx.x;
    if (c >= V.ncol) return;
    const int mi = V.mi, mj = V.mj, mk = V.mk, isles = G.isles;
    const int i = c % mi + 1, j = c / mi + 1;
    const size_t nub = (size_t)2 * (mi + 2) * (mj + 1);
    double ubc[2];
    for (int l = 1; l <= 2; l++) {
        const size_t o = (l - 1) + 2 * ((size_t)i + (size_t)(mi + 2) * j);
        double sum = 0.0;
        for (int isl = 1; isl <= isles; isl++) sum = sum + G.ubisl[nub * (isl - 1) + o] * V.psibc(isl);
        ubc[l - 1] = isles > 0 ? V.W[V.ML.ub + o] + sum : V.W[V.ML.ub + o];
    }
    const double s0 = V.csum(1, i, j), s1 = V.csum(2, i, j), rel = G.rel;
    for (int k = V.K1(i, j); k <= mk; k++) {
        if (V.K1(i + 1, j) <= k) {
            double y = V.uu(1, i, j, k) - s0 * V.rh(1, i, j) + ubc[0];
            y = rel * V.u1(1, i, j, k) + (1.0 - rel) * y;
            V.uu(1, i, j, k) = y;
            V.u1(1, i, j, k) = y;
        }
        if (V.K1(i, j + 1) <= k) {
            double y = V.uu(2, i, j, k) - s1 * V.rh(2, i, j) + ubc[1];
            y = rel * V.u1(2, i, j, k) + (1.0 - rel) * y;
            V.uu(2, i, j, k) = y;
            V.u1(2, i, j, k) = y;
        }
    }
}

// w from continuity, integrated up from the bottom (3661-3678); for an FP32
// member then the six face coefficients of the flux (with the
// Peclet-weighted upwinding of goldstein.f90:2489-2523) and the discrete
// divergence of the open faces, in FP64 from the FP64 velocities, rounded
// once -- the arithmetic of cgo_set_velocity.  Then the island correction
// of the stored ub and psi (219-233), which nothing in the step reads again.
__global__ __launch_bounds__(MOM_NT) void k_mom_cols_c(const MemberDev* __restrict__ mems,
                                                       const int32_t* __restrict__ list) {
    const MomView V(mems[list[blockIdx.y]]);
    const MomGridDev& G = V.G;
    const int x = blockIdx.x * MOM_NT + threadIdx.x;
    const int mi = V.mi, mj = V.mj, mk = V.mk, ncol = V.ncol, ncell = V.ncell, isles = G.isles;
    if (isles > 0) {
        const size_t nub = (size_t)2 * (mi + 2) * (mj + 1), nps = (size_t)(mi + 1) * (mj + 1);
        if (x < 2 * (mi + 2) * mj) {
            const size_t o = (size_t)(x % 2) + 2 * ((size_t)((x / 2) % (mi + 2)) + (size_t)(mi + 2) * (x / (2 * (mi + 2)) + 1));
            double sum = 0.0;
            for (int isl = 1; isl <= isles; isl++) sum = sum + G.ubisl[nub * (isl - 1) + o] * V.psibc(isl);
            V.W[V.ML.ub + o] = V.W[V.ML.ub + o] + sum;
        }
        if (x < (int)nps) {
            double sum = 0.0;
            for (int isl = 1; isl <= isles; isl++) sum = sum + G.psisl[nps * (isl - 1) + x] * V.psibc(isl);
            V.W[V.ML.psi + x] = V.W[V.ML.psi + x] + sum;
        }
    }
    if (x >= ncol) return;
    const int cc = x, i = cc % mi + 1, j = cc / mi + 1, kb = V.K1(i, j);
    double tv = 0;
    for (int k = kb; k <= mk - 1; k++) {
        const double tv1 = (V.uu(1, i, j, k) - V.uu(1, i == 1 ? mi : i - 1, j, k)) * G.rdphi * V.jt(JT_RC, j);
        const double vs = j > 1 ? V.uu(2, i, j - 1, k) : 0.0;
        const double tv2 = (V.uu(2, i, j, k) * V.jt(JT_CV, j) - vs * V.jt(JT_CV, j - 1)) * V.jt(JT_RDS, j);
        const double w = tv - G.dz[k] * (tv1 + tv2);
        V.uu(3, i, j, k) = w;
        tv = w;
    }
    const MemberDev& M = V.M;
    if (M.prec == 0) return;
    float* __restrict__ F = (float*)M.u;
    const double* __restrict__ U = V.U;
    const GridDev& g = M.g;
    const int JS = mj + 2, KS = mk + 2, I = i, J = j;
    const double diff1 = M.diff1, diffv = M.diff2;
    auto upwind = [](double pec) { return pec / (2.0 + fabs(pec)); };
    for (int k = kb; k <= mk; k++) {
        const int k1c = kb;
        const size_t o = (size_t)(k - 1) * ncol;
        const size_t c = o + cc;
        const size_t cw = o + (size_t)mi * (J - 1) + (I == 1 ? mi - 1 : I - 2);
        const size_t cs = o + (size_t)mi * (J > 1 ? J - 2 : 0) + (I - 1);
        const bool oe = k >= max(k1c, V.K1(I + 1, J)), ow = k >= max(V.K1(I - 1, J), k1c);
        const bool on = k >= max(k1c, V.K1(I, J + 1)), os = k >= max(V.K1(I, J - 1), k1c);
        const double rc = g.mj[MJ_RC * JS + J], rds = g.mj[MJ_RDS * JS + J];
        const double cvn = g.mj[MJ_CV * JS + J], cvs = g.mj[MJ_CV * JS + J - 1];
        const double rdz = g.mk[MK_RDZ * KS + k];
        const double ue = U[c], uw = U[cw], un = U[ncell + c], us = U[ncell + cs];
        const double ua = k == mk ? 0.0 : U[2 * (size_t)ncell + c];
        const double wb = k > k1c ? U[2 * (size_t)ncell + c - ncol] : 0.0;
        const double div = ((oe ? ue * rc : 0.0) - (ow ? uw * rc : 0.0)) * g.rdphi +
                           ((on ? cvn * un : 0.0) - (os ? cvs * us : 0.0)) * rds + (ua - wb) * rdz;
        const double rc2 = g.mj[MJ_RC2 * JS + J];
        const double cv2n = g.mj[MJ_CV2 * JS + J], cv2s = g.mj[MJ_CV2 * JS + J - 1];
        const double dsvn = g.mj[MJ_DSVM * JS + J], dsvs = g.mj[MJ_DSVM * JS + J - 1];
        const double dzak = g.mk[MK_DZA * KS + k], rdza = g.mk[MK_RDZA * KS + k];
        const double upe = upwind(ue * g.dphi / diff1), upw = upwind(uw * g.dphi / diff1);
        const double upn = upwind(un * dsvn / diff1), ups = upwind(us * dsvs / diff1);
        const double upa = upwind(ua * dzak / diffv);
        const bool top = k == mk;
        F[c] = oe ? (float)((ue * rc * (1.0 - upe) * 0.5 - rc2 * diff1) * g.rdphi) : 0.0f;
        F[ncell + c] = ow ? (float)((uw * rc * (1.0 + upw) * 0.5 + rc2 * diff1) * g.rdphi) : 0.0f;
        F[2 * (size_t)ncell + c] = on ? (float)((cvn * un * (1.0 - upn) * 0.5 - cv2n * diff1) * rds) : 0.0f;
        F[3 * (size_t)ncell + c] = os ? (float)((cvs * us * (1.0 + ups) * 0.5 + cv2s * diff1) * rds) : 0.0f;
        F[4 * (size_t)ncell + c] = top ? 0.0f : (float)(ua * (1.0 - upa) * 0.5 - rdza * diffv);
        F[5 * (size_t)ncell + c] = top ? 0.0f : (float)(-(ua * (1.0 + upa) * 0.5) - rdza * diffv);
        F[6 * (size_t)ncell + c] = (float)div;
    }
}

// LDS of the solve block of a member with N psi points and band n
size_t mom_solve_lds(int N, int n) {
    return (size_t)(N + MOM_MAXISL + 128 + (size_t)MOM_RING * MOM_CH * (n + 3)) * sizeof(double);
}
// ------------------------------------------------------------ the dense solve
// psi = A^-1 g for the FP32 members of one grid at once: C(i, c) =
// sum_k A(i, k) B(k, c), B's column c the staged forcing of the group's
// member c and C's column c its psi.  64 x 64 tiles of C per workgroup, k in
// steps of 16 through LDS with the next step's loads issued before the
// current one's products; each wave forms 16 rows x 64 members as four 16 x
// 16 tiles on the FP64 matrix cores (v_mfma_f64_16x16x4f64: lane i + 16k
// holds A(i, k) and B(k, j = i); item r of lane l is C(l / 16 + 4 r, l % 16)).
// The k range is cut into MOM_SPLITK slices (blockIdx.z), each writing its
// partial sums, which the solve block adds in slice order.
constexpr int GT = 64, GK = 16;
__global__ __launch_bounds__(256) void k_mom_gemm(const MemberDev* __restrict__ mems, const int32_t* __restrict__ list,
                                                   int nm) {
    __shared__ double As[GK][GT + 1];
    __shared__ double Bs[GK][GT + 1];
    const MomGridDev& G = *mems[list[0]].mg;
    const int N = G.N;
    const MomMemberLayout L = mom_member_layout(G.maxi, G.maxj, G.maxk, true);
    const double* __restrict__ A = G.ainv;
    const int r0 = blockIdx.x * GT, c0 = blockIdx.y * GT;
    const int slice = (N + MOM_SPLITK - 1) / MOM_SPLITK;
    const int kb = blockIdx.z * slice, ke = min(N, kb + slice);
    const int tx = threadIdx.x % 16, ty = threadIdx.x / 16;
    // each thread stages 4 values of A and 4 of B per step: rows / members
    // rr = ty + 16 q and k offset tx, so its members are fixed
    const double* bcol[4];
#pragma unroll
    for (int q = 0; q < 4; q++) {
        const int c = c0 + ty + 16 * q;
        bcol[q] = c < nm ? mems[list[c]].mw + L.gst : nullptr;
    }
    typedef double d4 __attribute__((ext_vector_type(4)));
    const int wv = threadIdx.x / 64, lane = threadIdx.x % 64, li = lane % 16, lk = lane / 16;
    d4 acc[4];
#pragma unroll
    for (int t = 0; t < 4; t++) acc[t] = d4{0.0, 0.0, 0.0, 0.0};
    double na[4], nb[4];
    auto fetch = [&](int k0) {
        const int k = k0 + tx;
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = r0 + ty + 16 * q;
            na[q] = (r < N && k < ke) ? A[(size_t)r * N + k] : 0.0;
            nb[q] = (bcol[q] && k < ke) ? bcol[q][k] : 0.0;
        }
    };
    fetch(kb);
    for (int k0 = kb; k0 < ke; k0 += GK) {
#pragma unroll
        for (int q = 0; q < 4; q++) {
            As[tx][ty + 16 * q] = na[q];
            Bs[tx][ty + 16 * q] = nb[q];
        }
        __syncthreads();
        if (k0 + GK < ke) fetch(k0 + GK);
#pragma unroll
        for (int ks = 0; ks < GK / 4; ks++) {
            const double a = As[ks * 4 + lk][wv * 16 + li];
#pragma unroll
            for (int t = 0; t < 4; t++)
                acc[t] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, Bs[ks * 4 + lk][t * 16 + li], acc[t], 0, 0, 0);
        }
        __syncthreads();
    }
#pragma unroll
    for (int t = 0; t < 4; t++) {
        const int c = c0 + t * 16 + li;
        if (c >= nm) continue;
        double* out = mems[list[c]].mw + L.gst + (size_t)N * (1 + blockIdx.z);
#pragma unroll
        for (int q = 0; q < 4; q++) {
            const int r = r0 + wv * 16 + lk + 4 * q;
            if (r < N) out[r] = acc[t][q];
        }
    }
}

// island path terms the solve block's ring holds
size_t mom_ring_terms(int n) { return (size_t)MOM_RING * MOM_CH * (n + 3); }

// part 1: the column kernel before the solve; 2: the rest; 3: both.
// list: every member with a velocity solve (the column kernels); band: those
// solved through the band (FP64 members, FP32 members of grids without the
// inverse); dense: the FP32 members of grids with the inverse, grouped by
// grid, groups[g] = {offset into dense, count, N}
void launch_mom(const MemberDev* mems, const int32_t* list, int nlist, const int32_t* band, int nband,
                const int32_t* dense, const int* groups, int ngroups, int maxN, int maxcol, int maxn, int part,
                hipStream_t s) {
    if (nlist <= 0) return;
    const int cb = (maxcol + MOM_NT - 1) / MOM_NT;
    // the stored ub and psi have at most (maxi+2) maxj x 2 and (maxi+1)(maxj+1)
    // entries: fewer than 4 x the columns + 256
    const int cb2 = (4 * maxcol + 256 + MOM_NT - 1) / MOM_NT;
    const size_t lds = mom_solve_lds(maxN, maxn);
    if (part & 1) hipLaunchKernelGGL(k_mom_cols_a, dim3(cb, nlist), dim3(MOM_NT), 0, s, mems, list);
    if (!(part & 2)) return;
    if (nband > 0) hipLaunchKernelGGL(k_mom_solve, dim3(nband), dim3(MOM_NT), lds, s, mems, band, 0);
    int ndense = 0;
    for (int q = 0; q < ngroups; q++) ndense += groups[3 * q + 1];
    if (ndense > 0) {
        hipLaunchKernelGGL(k_mom_solve, dim3(ndense), dim3(MOM_NT), lds, s, mems, dense, 1);
        for (int q = 0; q < ngroups; q++) {
            const int off = groups[3 * q], nm = groups[3 * q + 1], N = groups[3 * q + 2];
            hipLaunchKernelGGL(k_mom_gemm, dim3((N + GT - 1) / GT, (nm + GT - 1) / GT, MOM_SPLITK), dim3(256), 0, s,
                               mems, dense + off, nm);
        }
        hipLaunchKernelGGL(k_mom_solve, dim3(ndense), dim3(MOM_NT), lds, s, mems, dense, 2);
    }
    hipLaunchKernelGGL(k_mom_cols_b, dim3(cb, nlist), dim3(MOM_NT), 0, s, mems, list);
    hipLaunchKernelGGL(k_mom_cols_c, dim3(cb2, nlist), dim3(MOM_NT), 0, s, mems, list);
}

}  // namespace cgo
