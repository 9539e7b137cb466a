// k_mom.hip -- GOLDSTEIN's velocity solve on the device: the barotropic
// streamfunction with islands and the baroclinic velocities of every member
// with a wind stress (goldstein.f90:198-233: wind 3685-3714, jbar 3217-3315,
// ubarsolv 3500-3565, the island correction with island() of
// goldstein_lib.f90:186-241 and matmult 3470-3492, velc 3568-3679).
//
// One workgroup of 256 threads per member.  The column work (bottom pressure,
// the psi forcing, the shear integral, the barotropic correction, w) is spread
// over the threads; the banded solve is a chain of dependent operations
// (elimination step i needs row i final; back substitution subtracts the
// terms of a row in the reference's order) and runs on one wave, with the
// psi forcing in LDS and the factor rows streamed from L2 a few rows ahead,
// while the other three waves integrate the baroclinic shear, which does not
// depend on it.  Every operation rounds as written (-ffp-contract=off) in
// the reference's order, so an FP64 member's velocities are the reference
// arithmetic's.  For an FP32 member the kernel then forms the seven flux
// coefficient planes from the FP64 velocities, as cgo_set_velocity does.
#include "cgocn_dev.h"

namespace cgo {

constexpr int MOM_NT = 256;
constexpr int MOM_PF = 4;  // factor rows loaded ahead of the elimination and the back substitution

__device__ __forceinline__ double readlane_d(double v, int lane) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
    return __hiloint2double(hi, lo);
}

__global__ __launch_bounds__(MOM_NT) void k_mom(const MemberDev* __restrict__ mems, const int32_t* __restrict__ list) {
    extern __shared__ double lds[];
    const MemberDev& M = mems[list[blockIdx.x]];
    const MomGridDev& G = *M.mg;
    const int mi = G.maxi, mj = G.maxj, mk = G.maxk, n = G.n, N = G.N, isles = G.isles;
    const int ncol = mi * mj, ncell = ncol * mk;
    const int tid = threadIdx.x;
    double* gbs = lds;             // the psi forcing, then psi (N)
    double* erhs = lds + N;        // island integrals, then psibc (isles)

    const MomMemberLayout ML = mom_member_layout(mi, mj, mk, M.prec != 0);
    double* __restrict__ W = M.mw;
    double* __restrict__ U = M.u64;

    auto K1 = [&](int i, int j) { return k1at(M.g, i, j); };
    auto jt = [&](int t, int j) { return G.jt[t * (mj + 2) + j]; };
    auto rh = [&](int l, int i, int j) { return G.rh[(l - 1) + 3 * (i + (mi + 2) * j)]; };
    auto drag = [&](int l, int i, int j) { return G.drag[(l - 1) + 2 * (i + (mi + 2) * j)]; };
    auto ku = [&](int l, int i, int j) { return G.ku[(l - 1) + 2 * ((i - 1) + mi * (j - 1))]; };
    auto mkk = [&](int i, int j) { return G.mk[i + (mi + 2) * j]; };
    auto col = [&](int i, int j) { return (i - 1) + mi * (j - 1); };
    // rho(i, j, k) at a wet cell, i = 0..maxi+1 periodic
    auto rho = [&](int i, int j, int k) {
        const int ii = i == 0 ? mi : (i == mi + 1 ? 1 : i);
        return M.rho[(size_t)(k - 1) * ncol + col(ii, j)];
    };
    auto bp = [&](int i, int j, int k) -> double& { return W[ML.bp + i + (mi + 2) * (j + (mj + 2) * k)]; };
    auto sbp = [&](int i, int j) -> double& { return W[ML.sbp + i + (mi + 2) * j]; };
    auto ub = [&](int l, int i, int j) -> double& { return W[ML.ub + (l - 1) + 2 * (i + (mi + 2) * j)]; };
    auto psi = [&](int i, int j) -> double& { return W[ML.psi + i + (mi + 1) * j]; };
    auto tau = [&](int l, int i, int j) { return W[ML.tau + (l - 1) + 2 * col(i, j)]; };
    auto dztau = [&](int l, int i, int j) { return W[ML.dztau + (l - 1) + 2 * col(i, j)]; };
    auto dztav = [&](int l, int i, int j) { return W[ML.dztav + (l - 1) + 2 * col(i, j)]; };
    // u, v, w at interior points (i = 1..maxi, j = 1..maxj)
    auto uu = [&](int l, int i, int j, int k) -> double& { return U[(size_t)(l - 1) * ncell + (size_t)(k - 1) * ncol + col(i, j)]; };
    auto u1 = [&](int l, int i, int j, int k) -> double& { return W[ML.u1 + (size_t)(l - 1) * ncell + (size_t)(k - 1) * ncol + col(i, j)]; };
    auto csum = [&](int l, int i, int j) -> double& { return W[ML.csum + (size_t)(l - 1) * ncol + col(i, j)]; };
    const double dphi = G.dphi, rdphi = G.rdphi;

    // ---- jbar (3225-3248): bottom pressure from the bottom cell up, and its
    // integral above the level mk of the deepest wet neighbour
    for (int c = tid; c < ncol; c += MOM_NT) {
        const int i = c % mi + 1, j = c / mi + 1, kb = K1(i, j);
        if (kb > mk) continue;
        double b = bp(i, j, kb);  // never written: zero
        for (int k = kb + 1; k <= mk; k++) {
            b = b - (rho(i, j, k) + rho(i, j, k - 1)) * G.dza[k - 1] * 0.5;
            bp(i, j, k) = b;
            if (i == 1 && kb < mk) bp(mi + 1, j, k) = b;  // periodic copy (3249-3255)
        }
        if (mkk(i, j) > 0) {
            double s = 0;
            for (int k = mkk(i, j) + 1; k <= mk; k++) s = s + bp(i, j, k) * G.dz[k];
            sbp(i, j) = s;
        }
        if (i == 1) sbp(mi + 1, j) = sbp(1, j);
    }
    __syncthreads();

    // ---- the psi forcing: the wind part (wind) plus the torque of the
    // bottom pressure where the four columns round a psi point differ in depth
    // (jbar 3256-3313)
    for (int l = tid + 1; l <= N; l += MOM_NT) {
        const int i = (l - 1) % n + 1, j = (l - 1) / n;
        double g = W[ML.gbw + l - 1];
        if (j >= 1 && j <= mj - 1 && G.getj[col(i, j)]) {
            const int ip1 = i % mi + 1;
            auto span = [&](int k0, int k1_, int ii, int jj) {
                double t = 0;
                for (int k = k0; k <= k1_; k++) t = t + bp(ii, jj, k) * G.dz[k];
                return t;
            };
            double tv1 = span(ku(2, ip1, j), mkk(ip1, j + 1), ip1, j + 1);
            double tv2 = span(ku(2, ip1, j), mkk(ip1, j), ip1, j);
            double tv3 = span(ku(2, i, j), mkk(i, j + 1), i, j + 1);
            double tv4 = span(ku(2, i, j), mkk(i, j), i, j);
            g = g + ((tv3 + sbp(i, j + 1) - tv4 - sbp(i, j)) * rh(2, i, j) -
                     (tv1 + sbp(ip1, j + 1) - tv2 - sbp(ip1, j)) * rh(2, ip1, j)) *
                        rdphi * jt(JT_RDSV, j);
            tv1 = span(ku(1, i, j + 1), mkk(ip1, j + 1), ip1, j + 1);
            tv2 = span(ku(1, i, j), mkk(ip1, j), ip1, j);
            tv3 = span(ku(1, i, j + 1), mkk(i, j + 1), i, j + 1);
            tv4 = span(ku(1, i, j), mkk(i, j), i, j);
            g = g + ((tv1 + sbp(ip1, j + 1) - tv3 - sbp(i, j + 1)) * rh(1, i, j + 1) -
                     (tv2 + sbp(ip1, j) - tv4 - sbp(i, j)) * rh(1, i, j)) *
                        rdphi * jt(JT_RDSV, j);
        }
        gbs[l - 1] = g;
    }
    __syncthreads();

    if (tid < 64) {
        // ---- ubarsolv (3508-3522) on wave 0: elimination step i takes row i
        // (final) times the multiplier ratm(i+q, q) from row i+q, lane q-1
        const int lane = tid, q = lane + 1;
        const bool act = q <= n + 1;
        double rc[MOM_PF], rn[MOM_PF];
#pragma unroll
        for (int s = 0; s < MOM_PF; s++) {
            const int i = 1 + s;
            rc[s] = (act && i <= N - 1) ? G.fwd[(size_t)(i - 1) * (n + 1) + q - 1] : 0.0;
        }
        for (int i0 = 1; i0 <= N - 1; i0 += MOM_PF) {
#pragma unroll
            for (int s = 0; s < MOM_PF; s++) {
                const int i = i0 + MOM_PF + s;
                rn[s] = (act && i <= N - 1) ? G.fwd[(size_t)(i - 1) * (n + 1) + q - 1] : 0.0;
            }
#pragma unroll
            for (int s = 0; s < MOM_PF; s++) {
                const int i = i0 + s;
                if (i <= N - 1) {
                    const double x = gbs[i - 1];
                    if (act && i + q <= N) gbs[i + q - 1] = gbs[i + q - 1] - rc[s] * x;
                }
            }
#pragma unroll
            for (int s = 0; s < MOM_PF; s++) rc[s] = rn[s];
        }
        // back substitution: row i subtracts gap(i, n+2+k) G(i+k) for k = 1, 2, ..
        // in that order, then divides by the diagonal; lane k-1 forms the
        // product of term k >= 2 (rows solved earlier), the wave subtracts
        // them in turn after the term of the row just solved
        const int W2 = n + 2;
        double tp = gbs[N - 1] / G.bwd[(size_t)(N - 1) * W2];
        if (lane == 0) gbs[N - 1] = tp;
        double bc[MOM_PF], b0c[MOM_PF], b1c[MOM_PF], bn[MOM_PF], b0n[MOM_PF], b1n[MOM_PF];
        const int kl = lane + 1;
#pragma unroll
        for (int s = 0; s < MOM_PF; s++) {
            const int i = N - 1 - s;
            const bool ok = i >= 1;
            bc[s] = (ok && kl >= 2 && kl <= n + 1) ? G.bwd[(size_t)(i - 1) * W2 + kl] : 0.0;
            b0c[s] = ok ? G.bwd[(size_t)(i - 1) * W2] : 1.0;
            b1c[s] = ok ? G.bwd[(size_t)(i - 1) * W2 + 1] : 0.0;
        }
        for (int i0 = N - 1; i0 >= 1; i0 -= MOM_PF) {
#pragma unroll
            for (int s = 0; s < MOM_PF; s++) {
                const int i = i0 - MOM_PF - s;
                const bool ok = i >= 1;
                bn[s] = (ok && kl >= 2 && kl <= n + 1) ? G.bwd[(size_t)(i - 1) * W2 + kl] : 0.0;
                b0n[s] = ok ? G.bwd[(size_t)(i - 1) * W2] : 1.0;
                b1n[s] = ok ? G.bwd[(size_t)(i - 1) * W2 + 1] : 0.0;
            }
#pragma unroll
            for (int s = 0; s < MOM_PF; s++) {
                const int i = i0 - s;
                if (i >= 1) {
                    const int km = min(n + 1, N - i);
                    const double gk = gbs[min(i + kl - 1, N - 1)];
                    const double p = (kl >= 2 && kl <= km) ? bc[s] * gk : 0.0;
                    double t = gbs[i - 1] - b1c[s] * tp;
                    for (int k = 2; k <= km; k++) t = t - readlane_d(p, k - 1);
                    t = t / b0c[s];
                    if (lane == 0) gbs[i - 1] = t;
                    tp = t;
                }
            }
#pragma unroll
            for (int s = 0; s < MOM_PF; s++) {
                bc[s] = bn[s];
                b0c[s] = b0n[s];
                b1c[s] = b1n[s];
            }
        }
    } else {
        // ---- velc (3574-3640) on the other waves: the thermal-wind shear
        // (and the wind's Ekman shear in the top layer) integrated up each
        // column from zero at the bottom, and its depth sum
        const double* dza = G.dza;
        for (int c = tid - 64; c < ncol; c += MOM_NT - 64) {
            const int i = c % mi + 1, j = c / mi + 1;
            double sum0 = 0, sum1 = 0, dp0 = 0, dp1 = 0, up0 = 0, up1 = 0;
            const double sj = jt(JT_S, j), svj = jt(JT_SV, j), cj = jt(JT_C, j), cvj = jt(JT_CV, j);
            const double rtvij = G.rtv[col(i, j)], rtv3ij = G.rtv3[col(i, j)];
            const double d1 = drag(1, i, j), d2 = drag(2, i, j);
            for (int k = K1(i, j); k <= mk; k++) {
                double tv1, tv2, tv4, tv5;
                if (K1(i + 1, j) > k) {
                    tv1 = 0;
                    tv2 = 0;
                } else {
                    tv2 = -(rho(i + 1, j, k) - rho(i, j, k)) * rdphi * jt(JT_RC, j);
                    if (max(max(K1(i, j - 1), K1(i, j + 1)), max(K1(i + 1, j - 1), K1(i + 1, j + 1))) <= k)
                        tv1 = -cj * (rho(i + 1, j + 1, k) - rho(i + 1, j - 1, k) + rho(i, j + 1, k) - rho(i, j - 1, k)) *
                              jt(JT_RDS2, j) * 0.25;
                    else if (max(K1(i, j - 1), K1(i + 1, j - 1)) <= k)
                        tv1 = -cj * (rho(i + 1, j, k) - rho(i + 1, j - 1, k) + rho(i, j, k) - rho(i, j - 1, k)) *
                              jt(JT_RDSV, j - 1) * 0.5;
                    else if (max(K1(i, j + 1), K1(i + 1, j + 1)) <= k)
                        tv1 = -cj * (rho(i + 1, j + 1, k) - rho(i + 1, j, k) + rho(i, j + 1, k) - rho(i, j, k)) *
                              jt(JT_RDSV, j) * 0.5;
                    else
                        tv1 = 0;
                }
                if (K1(i, j + 1) > k) {
                    tv4 = 0;
                    tv5 = 0;
                } else {
                    tv4 = -cvj * (rho(i, j + 1, k) - rho(i, j, k)) * jt(JT_RDSV, j);
                    if (max(max(K1(i - 1, j), K1(i - 1, j + 1)), max(K1(i + 1, j), K1(i + 1, j + 1))) <= k)
                        tv5 = -(rho(i + 1, j + 1, k) - rho(i - 1, j + 1, k) + rho(i + 1, j, k) - rho(i - 1, j, k)) * rdphi *
                              0.25 * jt(JT_RCV, j);
                    else if (max(K1(i - 1, j), K1(i - 1, j + 1)) <= k)
                        tv5 = -(rho(i, j + 1, k) - rho(i - 1, j + 1, k) + rho(i, j, k) - rho(i - 1, j, k)) * rdphi * 0.5 *
                              jt(JT_RCV, j);
                    else if (max(K1(i + 1, j), K1(i + 1, j + 1)) <= k)
                        tv5 = -(rho(i + 1, j + 1, k) - rho(i, j + 1, k) + rho(i + 1, j, k) - rho(i, j, k)) * rdphi * 0.5 *
                              jt(JT_RCV, j);
                    else
                        tv5 = 0;
                }
                if (k == mk) {
                    if (K1(i + 1, j) <= k) {
                        tv1 = tv1 - dztau(2, i, j);
                        tv2 = tv2 - dztau(1, i, j);
                    }
                    if (K1(i, j + 1) <= k) {
                        tv4 = tv4 - dztav(2, i, j);
                        tv5 = tv5 - dztav(1, i, j);
                    }
                }
                const double dz0 = -(sj * tv1 + d1 * tv2) * rtvij;
                const double dz1 = -(d2 * tv4 - svj * tv5) * rtv3ij;
                double x0, x1;
                if (k == K1(i, j)) {
                    x0 = 0;
                    x1 = 0;
                } else {
                    x0 = up0 + dza[k - 1] * (dz0 + dp0) * 0.5;
                    sum0 = sum0 + G.dz[k] * x0;
                    x1 = up1 + dza[k - 1] * (dz1 + dp1) * 0.5;
                    sum1 = sum1 + G.dz[k] * x1;
                }
                uu(1, i, j, k) = x0;
                uu(2, i, j, k) = x1;
                up0 = x0;
                up1 = x1;
                dp0 = dz0;
                dp1 = dz1;
            }
            csum(1, i, j) = sum0;
            csum(2, i, j) = sum1;
        }
    }
    __syncthreads();

    // ---- psi and the barotropic velocities from the solution (3524-3565),
    // each boundary copy written by the thread that forms its source
    for (int l = tid + 1; l <= N; l += MOM_NT) {
        const int i = (l - 1) % n + 1, j = (l - 1) / n;
        const double p = gbs[l - 1];
        psi(i, j) = p;
        if (i == mi) psi(0, j) = p;
        if (j >= 1) {
            const double x = -rh(1, i, j) * jt(JT_C, j) * (p - gbs[l - 1 - n]) * jt(JT_RDS, j);
            ub(1, i, j) = x;
            if (i == mi) ub(1, 0, j) = x;
            if (i == 1) ub(1, mi + 1, j) = x;
        }
        double y;
        if (j >= 1 && j <= mj - 1) {
            const double pw = gbs[(i == 1 ? mi : i - 1) + j * n - 1];
            y = rh(2, i, j) * (p - pw) * jt(JT_RCV, j) * rdphi;
        } else {
            y = 0.0;
        }
        ub(2, i, j) = y;
        if (i == 1) ub(2, mi + 1, j) = y;
        if (i == mi) ub(2, 0, j) = y;
    }
    __syncthreads();

    // ---- the path integral round every island of the momentum balance with
    // wind and bottom-pressure torque (island, indj = 1), then the island
    // streamfunction values that cancel it (209-218)
    if (tid < isles) {
        const int isl = tid + 1;
        double e = 0.0;
        for (int p = 1; p <= G.npi[isl - 1]; p++) {
            const int q = (p - 1) + G.mpi * (isl - 1);
            const int lpi = G.lpisl[q], ipi = G.ipisl[q], jpi = G.jpisl[q];
            const int al = abs(lpi);
            double cor;
            if (al == 1)
                cor = -jt(JT_S, jpi) * 0.25 * (ub(2, ipi, jpi) + ub(2, ipi + 1, jpi) + ub(2, ipi, jpi - 1) + ub(2, ipi + 1, jpi - 1));
            else
                cor = jt(JT_SV, jpi) * 0.25 * (ub(1, ipi - 1, jpi) + ub(1, ipi, jpi) + ub(1, ipi - 1, jpi + 1) + ub(1, ipi, jpi + 1));
            const double sgn = lpi >= 0 ? 1.0 : -1.0;
            const double rhv = rh(al, ipi, jpi), tauv = tau(al, ipi, jpi);
            e = e + sgn * (drag(al, ipi, jpi) * ub(al, ipi, jpi) + cor - 1.0 * tauv * rhv) *
                        (jt(JT_C, jpi) * dphi * (2.0 - al) + jt(JT_RCV, jpi) * jt(JT_DSV, jpi) * (al - 1.0));
            if (al == 1) {
                double t = 0.0;
                for (int k = ku(1, ipi, jpi); k <= mkk(ipi + 1, jpi); k++) t = t + bp(ipi + 1, jpi, k) * G.dz[k];
                for (int k = ku(1, ipi, jpi); k <= mkk(ipi, jpi); k++) t = t - bp(ipi, jpi, k) * G.dz[k];
                e = e + (sbp(ipi + 1, jpi) - sbp(ipi, jpi) + t) * sgn * rh(1, ipi, jpi);
            } else {
                double t = 0.0;
                for (int k = ku(2, ipi, jpi); k <= mkk(ipi, jpi + 1); k++) t = t + bp(ipi, jpi + 1, k) * G.dz[k];
                for (int k = ku(2, ipi, jpi); k <= mkk(ipi, jpi); k++) t = t - bp(ipi, jpi, k) * G.dz[k];
                e = e + (sbp(ipi, jpi + 1) - sbp(ipi, jpi) + t) * sgn * rh(2, ipi, jpi);
            }
        }
        erhs[isl - 1] = e;
    }
    __syncthreads();
    if (tid == 0 && isles > 0) {
        auto A = [&](int a, int b) { return G.erisl[(a - 1) + isles * (b - 1)]; };
        if (isles > 1) {
            // matmult with the eliminated matrix of matinv_gold
            for (int i = 1; i <= isles - 1; i++)
                for (int j = i + 1; j <= isles; j++) erhs[j - 1] = A(i, i) * erhs[j - 1] - A(j, i) * erhs[i - 1];
            erhs[isles - 1] = erhs[isles - 1] / A(isles, isles);
            for (int i = isles - 1; i >= 1; i--) {
                for (int j = i + 1; j <= isles; j++) erhs[i - 1] = erhs[i - 1] - A(i, j) * erhs[j - 1];
                erhs[i - 1] = erhs[i - 1] / A(i, i);
            }
            for (int isl = 1; isl <= isles; isl++) erhs[isl - 1] = -erhs[isl - 1];
        } else {
            erhs[0] = -erhs[0] / A(1, 1);
        }
    }
    __syncthreads();

    // ---- add the island fields times their psi values (219-233)
    if (isles > 0) {
        const size_t nub = (size_t)2 * (mi + 2) * (mj + 1), nps = (size_t)(mi + 1) * (mj + 1);
        for (int x = tid; x < 2 * (mi + 2) * mj; x += MOM_NT) {
            const int l = x % 2 + 1, i = (x / 2) % (mi + 2), j = x / (2 * (mi + 2)) + 1;
            const size_t o = (l - 1) + 2 * ((size_t)i + (size_t)(mi + 2) * j);
            double sum = 0.0;
            for (int isl = 1; isl <= isles; isl++) sum = sum + G.ubisl[nub * (isl - 1) + o] * erhs[isl - 1];
            W[ML.ub + o] = W[ML.ub + o] + sum;
        }
        for (int x = tid; x < (mi + 1) * (mj + 1); x += MOM_NT) {
            double sum = 0.0;
            for (int isl = 1; isl <= isles; isl++) sum = sum + G.psisl[nps * (isl - 1) + x] * erhs[isl - 1];
            W[ML.psi + x] = W[ML.psi + x] + sum;
        }
    }
    __syncthreads();

    // ---- velc (3641-3657): the barotropic velocity in place of the shear's
    // depth mean at every open face, relaxed towards the last step's
    for (int c = tid; c < ncol; c += MOM_NT) {
        const int i = c % mi + 1, j = c / mi + 1;
        const double s0 = csum(1, i, j), s1 = csum(2, i, j);
        for (int k = K1(i, j); k <= mk; k++) {
            if (K1(i + 1, j) <= k) {
                double x = uu(1, i, j, k) - s0 * rh(1, i, j) + ub(1, i, j);
                x = G.rel * u1(1, i, j, k) + (1.0 - G.rel) * x;
                uu(1, i, j, k) = x;
                u1(1, i, j, k) = x;
            }
            if (K1(i, j + 1) <= k) {
                double x = uu(2, i, j, k) - s1 * rh(2, i, j) + ub(2, i, j);
                x = G.rel * u1(2, i, j, k) + (1.0 - G.rel) * x;
                uu(2, i, j, k) = x;
                u1(2, i, j, k) = x;
            }
        }
    }
    __syncthreads();

    // ---- w from continuity, integrated up from the bottom (3661-3678)
    for (int c = tid; c < ncol; c += MOM_NT) {
        const int i = c % mi + 1, j = c / mi + 1;
        double tv = 0;
        for (int k = K1(i, j); k <= mk - 1; k++) {
            const double tv1 = (uu(1, i, j, k) - uu(1, i == 1 ? mi : i - 1, j, k)) * rdphi * jt(JT_RC, j);
            const double vs = j > 1 ? uu(2, i, j - 1, k) : 0.0;
            const double tv2 = (uu(2, i, j, k) * jt(JT_CV, j) - vs * jt(JT_CV, j - 1)) * jt(JT_RDS, j);
            const double w = tv - G.dz[k] * (tv1 + tv2);
            uu(3, i, j, k) = w;
            tv = w;
        }
    }

    if (M.prec == 0) return;
    __syncthreads();

    // ---- FP32 form: the six face coefficients of the flux (with the
    // Peclet-weighted upwinding of goldstein.f90:2489-2523) and the discrete
    // divergence of the open faces, in FP64 from the FP64 velocities, rounded
    // once -- the arithmetic of cgo_set_velocity
    float* __restrict__ F = (float*)M.u;
    const GridDev& g = M.g;
    const int JS = mj + 2, KS = mk + 2;
    const double diff1 = M.diff1, diffv = M.diff2;
    auto upwind = [](double pec) { return pec / (2.0 + fabs(pec)); };
    for (int x = tid; x < ncell; x += MOM_NT) {
        const int k = x / ncol + 1, cc = x % ncol, I = cc % mi + 1, J = cc / mi + 1;
        const int k1c = K1(I, J);
        if (k < k1c) continue;
        const size_t o = (size_t)(k - 1) * ncol;
        const size_t c = o + cc;
        const size_t cw = o + (size_t)mi * (J - 1) + (I == 1 ? mi - 1 : I - 2);
        const size_t cs = o + (size_t)mi * (J > 1 ? J - 2 : 0) + (I - 1);
        const bool oe = k >= max(k1c, K1(I + 1, J)), ow = k >= max(K1(I - 1, J), k1c);
        const bool on = k >= max(k1c, K1(I, J + 1)), os = k >= max(K1(I, J - 1), k1c);
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

void launch_mom(const MemberDev* mems, const int32_t* list, int nlist, int maxN, hipStream_t s) {
    if (nlist <= 0) return;
    const size_t lds = (size_t)(maxN + MOM_MAXISL) * sizeof(double);
    hipLaunchKernelGGL(k_mom, dim3(nlist), dim3(MOM_NT), lds, s, mems, list);
}

}  // namespace cgo
