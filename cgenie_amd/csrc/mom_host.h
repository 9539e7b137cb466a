// mom_host.h -- host set-up of GOLDSTEIN's velocity solve for one grid: the
// seabed depths at T, u and v points, the drag, the psi equation in band form
// and its elimination, and the island basis (initialise_goldstein,
// goldstein.f90:1359-1414 and 1460-1575, 1788-1814; drgset 2845-2885;
// invert 3138-3204; ubarsolv 3500-3565 for the basis; island,
// goldstein_lib.f90:186-241; matinv_gold 3452-3467).
//
// Runs once per grid in FP64 with every operation rounded as written (the
// translation unit is compiled with -ffp-contract=off and x86-64 SSE2
// arithmetic), so the factors the device reads are the reference's own.
// Indices are the reference's (Fortran, with its lower bounds); arrays are
// flat with the layouts listed in MomGridDev (cgocn_dev.h).
#pragma once
#include <algorithm>
#include <cstdint>
#include <cstdlib>
#include <string>
#include <vector>

#include "cgocn_dev.h"

namespace cgo {

struct MomSetup {
    int mi = 0, mj = 0, mk = 0, n = 0, N = 0, isles = 0, mpi = 1;
    double dphi = 0, rdphi = 0;
    std::vector<int32_t> k1;                 // (0:maxi+1, 0:maxj+1)
    std::vector<double> jt;                  // [JT_N][maxj+2]
    std::vector<double> dz, dza;             // [maxk+2]
    std::vector<double> h, rh, drag, rtv, rtv3;
    std::vector<int32_t> ku, mkv;
    std::vector<uint8_t> getj;
    std::vector<double> gap, ratm;           // (N, 2n+3) and (N, n+1), rows 1-based
    std::vector<double> ubisl, psisl, erisl;
    std::vector<int32_t> npi, lpisl, ipisl, jpisl;

    int K1(int i, int j) const { return k1[(size_t)i + (size_t)(mi + 2) * j]; }
    double J(int t, int j) const { return jt[(size_t)t * (mj + 2) + j]; }
    size_t p2(int i, int j) const { return (size_t)i + (size_t)(mi + 2) * j; }
    double& H(int l, int i, int j) { return h[(l - 1) + 3 * p2(i, j)]; }
    double& RH(int l, int i, int j) { return rh[(l - 1) + 3 * p2(i, j)]; }
    double RH(int l, int i, int j) const { return rh[(l - 1) + 3 * p2(i, j)]; }
    double& DRAG(int l, int i, int j) { return drag[(l - 1) + 2 * p2(i, j)]; }
    double DRAG(int l, int i, int j) const { return drag[(l - 1) + 2 * p2(i, j)]; }
    size_t ij(int i, int j) const { return (size_t)(i - 1) + (size_t)mi * (j - 1); }
    double& GAP(int k, int l) { return gap[(size_t)(k - 1) * (2 * n + 3) + (l - 1)]; }
    double& RATM(int k, int l) { return ratm[(size_t)(k - 1) * (n + 1) + (l - 1)]; }
    static double& UB(double* u, int mi, int l, int i, int j) { return u[(l - 1) + 2 * ((size_t)i + (size_t)(mi + 2) * j)]; }
    static double& PS(double* p, int mi, int i, int j) { return p[(size_t)i + (size_t)(mi + 1) * j]; }
};

inline int isign1(int x) { return x >= 0 ? 1 : -1; }

// drgset: drag at psi points, raised by drgf over shallow water (a
// neighbourhood deeper index than kmxdrg) and in the equatorial band, then
// averaged onto the u and v points
inline void mom_drag(MomSetup& M, double adrag, double drgf, int kmxdrg, int jeb) {
    const int mi = M.mi, mj = M.mj;
    std::vector<double> t((size_t)(mi + 1) * (mj + 1), 0.0);
    auto T = [&](int i, int j) -> double& { return t[(size_t)i + (size_t)(mi + 1) * j]; };
    for (int j = 0; j <= mj; j++)
        for (int i = 0; i <= mi; i++) {
            const int near2 = std::max(std::max(M.K1(i, j), M.K1(i + 1, j)), std::max(M.K1(i, j + 1), M.K1(i + 1, j + 1)));
            int near4 = M.K1(i, j);
            for (int j1 = std::max(0, j - 1); j1 <= std::min(mj + 1, j + 2); j1++)
                for (int i1 = i - 1; i1 <= i + 2; i1++) near4 = std::max(near4, M.K1(1 + (mi + i1 - 1) % mi, j1));
            const int eq = std::abs(j - mj / 2);
            if (near2 > kmxdrg || eq <= jeb)
                T(i, j) = adrag * drgf * drgf;
            else if (near4 > kmxdrg || eq == jeb + 1)
                T(i, j) = adrag * drgf;
            else
                T(i, j) = adrag;
        }
    for (int j = 1; j <= mj; j++)
        for (int i = 1; i <= mi; i++) {
            M.DRAG(1, i, j) = 0.5 * (T(i, j) + T(i, j - 1));
            M.DRAG(2, i, j) = 0.5 * (T(i, j) + T(i - 1, j));
        }
    for (int j = 0; j <= mj + 1; j++) M.DRAG(2, mi + 1, j) = M.DRAG(2, 1, j);
}

// invert: the psi equation at psi point k = i + j maxi (i = 1..maxi periodic,
// j = 0..maxj) as a band of half-width maxi+1 centred on column n+2, then its
// elimination without pivoting, keeping the multipliers
inline void mom_band(MomSetup& M) {
    const int n = M.n, mi = M.mi, mj = M.mj, N = M.N;
    const double dphi = M.dphi, rdphi = M.rdphi;
    M.gap.assign((size_t)N * (2 * n + 3), 0.0);
    M.ratm.assign((size_t)N * (n + 1), 0.0);
    auto s = [&](int j) { return M.J(JT_S, j); };
    auto sv = [&](int j) { return M.J(JT_SV, j); };
    auto c = [&](int j) { return M.J(JT_C, j); };
    auto cv = [&](int j) { return M.J(JT_CV, j); };
    auto ds = [&](int j) { return M.J(JT_DS, j); };
    auto dsv = [&](int j) { return M.J(JT_DSV, j); };
    auto rcv = [&](int j) { return M.J(JT_RCV, j); };
    for (int i = 1; i <= mi; i++)
        for (int j = 0; j <= mj; j++) {
            const int k = i + j * n;
            const bool wet = std::max(std::max(M.K1(i, j), M.K1(i + 1, j)), std::max(M.K1(i, j + 1), M.K1(i + 1, j + 1))) <= M.mk;
            if (!wet) {
                M.GAP(k, n + 2) = 1;
                continue;
            }
            const double tv = (s(j + 1) * M.RH(1, i, j + 1) - s(j) * M.RH(1, i, j)) / (2.0 * dsv(j) * dphi);
            const double tv1 = (sv(j) * M.RH(2, i + 1, j) - sv(j) * M.RH(2, i, j)) / (2.0 * dphi * dsv(j));
            const double south = M.DRAG(1, i, j) * c(j) * c(j) * M.RH(1, i, j);
            const double north = M.DRAG(1, i, j + 1) * c(j + 1) * c(j + 1) * M.RH(1, i, j + 1);
            M.GAP(k, 2) = south / (ds(j) * dsv(j)) + tv1;
            M.GAP(k, i == 1 ? 2 * n + 1 : n + 1) = M.DRAG(2, i, j) * rcv(j) * rcv(j) * rdphi * rdphi * M.RH(2, i, j) - tv;
            M.GAP(k, n + 2) = -(M.DRAG(2, i, j) * M.RH(2, i, j) + M.DRAG(2, i + 1, j) * M.RH(2, i + 1, j)) /
                                  (cv(j) * cv(j) * dphi * dphi) -
                              (south / ds(j) + north / ds(j + 1)) / dsv(j);
            M.GAP(k, i == mi ? 3 : n + 3) = M.DRAG(2, i + 1, j) * M.RH(2, i + 1, j) / (cv(j) * cv(j) * dphi * dphi) + tv;
            M.GAP(k, 2 * n + 2) = north / (ds(j + 1) * dsv(j)) - tv1;
        }
    for (int i = 1; i <= N - 1; i++) {
        const int last = std::min(i + n + 1, N);
        for (int j = i + 1; j <= last; j++) {
            const double rat = M.GAP(j, n + 2 - j + i) / M.GAP(i, n + 2);
            M.RATM(j, j - i) = rat;
            if (rat != 0)
                for (int k = n + 2 - j + i; k <= 2 * n + 3 - j + i; k++) M.GAP(j, k) = M.GAP(j, k) - rat * M.GAP(i, k + j - i);
        }
    }
}

// ubarsolv on a right-hand side gb (1-based, N entries): the eliminated band's
// forward and back substitution, psi, and the barotropic velocities with
// their boundary copies
// The psi equation's inverse, column c being the band solve of the unit
// vector e_c through the same factors (the forward elimination leaves the
// rows above c at zero), row-major: ainv[i N + c].  For the FP32 form, whose
// members multiply their forcing by it instead of substituting through the
// band one row after another.
inline void mom_inverse(MomSetup& M, std::vector<double>& ainv) {
    const int n = M.n, N = M.N;
    ainv.assign((size_t)N * N, 0.0);
    std::vector<double> g(N + 1);
    for (int c = 1; c <= N; c++) {
        std::fill(g.begin(), g.end(), 0.0);
        g[c] = 1.0;
        for (int i = c; i <= N - 1; i++) {
            if (g[i] == 0.0) continue;
            const int last = std::min(i + n + 1, N);
            for (int j = i + 1; j <= last; j++) g[j] = g[j] - M.RATM(j, j - i) * g[i];
        }
        g[N] = g[N] / M.GAP(N, n + 2);
        for (int i = N - 1; i >= 1; i--) {
            const int km = std::min(n + 1, N - i);
            double v = g[i];
            for (int k = 1; k <= km; k++) v = v - M.GAP(i, n + 2 + k) * g[i + k];
            g[i] = v / M.GAP(i, n + 2);
        }
        for (int i = 1; i <= N; i++) ainv[(size_t)(i - 1) * N + (c - 1)] = g[i];
    }
}

inline void mom_solve(MomSetup& M, std::vector<double>& gb, double* ub, double* psi) {
    const int n = M.n, N = M.N, mi = M.mi, mj = M.mj;
    auto G = [&](int k) -> double& { return gb[k - 1]; };
    for (int i = 1; i <= N - 1; i++) {
        const int last = std::min(i + n + 1, N);
        for (int j = i + 1; j <= last; j++) G(j) = G(j) - M.RATM(j, j - i) * G(i);
    }
    G(N) = G(N) / M.GAP(N, n + 2);
    for (int i = N - 1; i >= 1; i--) {
        const int km = std::min(n + 1, N - i);
        for (int k = 1; k <= km; k++) G(i) = G(i) - M.GAP(i, n + 2 + k) * G(i + k);
        G(i) = G(i) / M.GAP(i, n + 2);
    }
    for (int j = 0; j <= mj; j++) {
        for (int i = 1; i <= mi; i++) MomSetup::PS(psi, mi, i, j) = G(i + j * n);
        MomSetup::PS(psi, mi, 0, j) = MomSetup::PS(psi, mi, mi, j);
    }
    auto U = [&](int l, int i, int j) -> double& { return MomSetup::UB(ub, mi, l, i, j); };
    auto P = [&](int i, int j) { return MomSetup::PS(psi, mi, i, j); };
    for (int j = 1; j <= mj; j++)
        for (int i = 1; i <= mi; i++) U(1, i, j) = -M.RH(1, i, j) * M.J(JT_C, j) * (P(i, j) - P(i, j - 1)) * M.J(JT_RDS, j);
    for (int j = 1; j <= mj - 1; j++)
        for (int i = 1; i <= mi; i++) U(2, i, j) = M.RH(2, i, j) * (P(i, j) - P(i - 1, j)) * M.J(JT_RCV, j) * M.rdphi;
    for (int i = 1; i <= mi; i++) {
        U(2, i, mj) = 0.0;
        U(2, i, 0) = 0.0;
    }
    for (int j = 1; j <= mj; j++) {
        U(2, mi + 1, j) = U(2, 1, j);
        U(1, 0, j) = U(1, mi, j);
        U(1, mi + 1, j) = U(1, 1, j);
        U(2, 0, j) = U(2, mi, j);
    }
    U(2, mi + 1, 0) = U(2, 1, 0);
    U(2, 0, 0) = U(2, mi, 0);
}

// island with indj = 0: the path integral round island isl of the drag and
// Coriolis terms of a barotropic velocity field
inline double mom_path_integral(const MomSetup& M, const double* ub, int isl) {
    const int mi = M.mi;
    auto U = [&](int l, int i, int j) { return ub[(l - 1) + 2 * ((size_t)i + (size_t)(mi + 2) * j)]; };
    double e = 0.0;
    for (int p = 1; p <= M.npi[isl - 1]; p++) {
        const size_t q = (size_t)(p - 1) + (size_t)M.mpi * (isl - 1);
        const int lpi = M.lpisl[q], ipi = M.ipisl[q], jpi = M.jpisl[q], al = std::abs(lpi);
        double cor;
        if (al == 1)
            cor = -M.J(JT_S, jpi) * 0.25 * (U(2, ipi, jpi) + U(2, ipi + 1, jpi) + U(2, ipi, jpi - 1) + U(2, ipi + 1, jpi - 1));
        else
            cor = M.J(JT_SV, jpi) * 0.25 * (U(1, ipi - 1, jpi) + U(1, ipi, jpi) + U(1, ipi - 1, jpi + 1) + U(1, ipi, jpi + 1));
        const double len = M.J(JT_C, jpi) * M.dphi * (2.0 - al) + M.J(JT_RCV, jpi) * M.J(JT_DSV, jpi) * (al - 1.0);
        // (the wind term, indj * tau * rh, is exactly zero here)
        e = e + (double)isign1(lpi) * (M.DRAG(al, ipi, jpi) * U(al, ipi, jpi) + cor) * len;
    }
    return e;
}

// The whole set-up; returns "" or what is wrong with the island data.
// labels (maxi, 0:maxj): the landmass number of every psi point (the .psiles
// file); paths (3, sum npi): lpi, ipi, jpi of every path point in order.
inline std::string mom_setup(MomSetup& M, const int32_t* k1, int mi, int mj, int mk, const double* const* jrows,
                             const double* dz, const double* dza, double dphi, double rdphi, const int32_t* labels,
                             int isles, const int32_t* npi, const int32_t* paths, double adrag, double fsc) {
    M.mi = mi;
    M.mj = mj;
    M.mk = mk;
    M.n = mi;
    M.N = mi * (mj + 1);
    M.isles = isles;
    M.dphi = dphi;
    M.rdphi = rdphi;
    const size_t P2 = (size_t)(mi + 2) * (mj + 2);
    M.k1.assign(k1, k1 + P2);
    M.jt.assign((size_t)JT_N * (mj + 2), 0.0);
    for (int t = 0; t < JT_N; t++)
        for (int j = 0; j < mj + 2; j++) M.jt[(size_t)t * (mj + 2) + j] = jrows[t][j];
    M.dz.assign(dz, dz + mk + 2);
    M.dza.assign(dza, dza + mk + 2);
    M.h.assign(3 * P2, 0.0);
    M.rh.assign(3 * P2, 0.0);
    M.drag.assign(2 * P2, 0.0);
    M.rtv.assign((size_t)mi * mj, 0.0);
    M.rtv3.assign((size_t)mi * mj, 0.0);
    M.ku.assign((size_t)2 * mi * mj, 0);
    M.mkv.assign(P2, 0);
    M.getj.assign((size_t)mi * mj, 0);
    // seabed depth and its reciprocal at T points, then the shallower of the
    // two neighbours at u and v points (1359-1384)
    for (int j = mj + 1; j >= 0; j--)
        for (int i = 0; i <= mi + 1; i++)
            if (M.K1(i, j) <= mk) {
                for (int k = M.K1(i, j); k <= mk; k++) M.H(3, i, j) = M.H(3, i, j) + dz[k];
                M.RH(3, i, j) = 1.0 / M.H(3, i, j);
            }
    for (int j = 0; j <= mj + 1; j++)
        for (int i = 0; i <= mi; i++) {
            M.H(1, i, j) = std::min(M.H(3, i, j), M.H(3, i + 1, j));
            if (std::max(M.K1(i, j), M.K1(i + 1, j)) <= mk) M.RH(1, i, j) = 1.0 / M.H(1, i, j);
        }
    for (int j = 0; j <= mj; j++)
        for (int i = 0; i <= mi + 1; i++) {
            M.H(2, i, j) = std::min(M.H(3, i, j), M.H(3, i, j + 1));
            if (std::max(M.K1(i, j), M.K1(i, j + 1)) <= mk) M.RH(2, i, j) = 1.0 / M.H(2, i, j);
        }
    for (int j = 1; j <= mj; j++)
        for (int i = 1; i <= mi; i++) {
            M.ku[2 * M.ij(i, j)] = std::max(M.K1(i, j), M.K1(i + 1, j));
            M.ku[2 * M.ij(i, j) + 1] = std::max(M.K1(i, j), M.K1(i, j + 1));
        }
    // drag (1395-1405) and the reciprocal Coriolis-drag denominators (1409-1414)
    mom_drag(M, 1.0 / (adrag * 86400 * fsc), 3.0, mk / 2, 1);
    for (int j = 1; j <= mj; j++)
        for (int i = 1; i <= mi; i++) {
            M.rtv[M.ij(i, j)] = 1.0 / (M.J(JT_S, j) * M.J(JT_S, j) + M.DRAG(1, i, j) * M.DRAG(1, i, j));
            M.rtv3[M.ij(i, j)] = 1.0 / (M.J(JT_SV, j) * M.J(JT_SV, j) + M.DRAG(2, i, j) * M.DRAG(2, i, j));
        }
    // mk: the deepest bottom index among a wet column and its four wet
    // neighbours (0 over land); getj: wet psi points whose four columns differ
    // in depth, where the JEBAR term is non-zero (1460-1493)
    auto wetf = [&](int i, int j) { return (1 + isign1(mk - M.K1(i, j))) / 2; };
    for (int j = 1; j <= mj; j++)
        for (int i = 1; i <= mi; i++) {
            const int v = std::max(std::max(M.K1(i, j) * wetf(i, j), M.K1(i + 1, j) * wetf(i + 1, j)),
                                   std::max(std::max(M.K1(i - 1, j) * wetf(i - 1, j), M.K1(i, j + 1) * wetf(i, j + 1)),
                                            M.K1(i, j - 1) * wetf(i, j - 1)));
            M.mkv[M.p2(i, j)] = v * wetf(i, j);
        }
    for (int j = 1; j <= mj; j++) M.mkv[M.p2(mi + 1, j)] = M.mkv[M.p2(1, j)];
    for (int j = 1; j <= mj; j++)
        for (int i = 1; i <= mi; i++) {
            const int a = M.K1(i, j), b = M.K1(i + 1, j), c = M.K1(i, j + 1), d = M.K1(i + 1, j + 1);
            M.getj[M.ij(i, j)] = std::max(std::max(a, b), std::max(c, d)) <= mk && (a != c || a != b || a != d);
        }
    // a wet psi point on the southern or northern boundary row has no wind
    // stress to force it (wind, 3697-3700)
    for (int i = 1; i <= mi; i++)
        for (int j : {0, mj})
            if (std::max(std::max(M.K1(i, j), M.K1(i + 1, j)), std::max(M.K1(i, j + 1), M.K1(i + 1, j + 1))) <= mk)
                return "psi point (" + std::to_string(i) + ", " + std::to_string(j) + ") on a boundary row is wet";
    // islands (1505-1575): paths checked as the reference checks them, plus
    // the index ranges island() reads: ku(2, maxi, maxj), ub(2, 0:maxi+1, 0:maxj)
    if (isles < 0 || isles > MOM_MAXISL) return "isles " + std::to_string(isles) + " outside 0.." + std::to_string(MOM_MAXISL);
    int mpi = 1;
    for (int q = 0; q < isles; q++) {
        if (npi[q] < 1) return "island " + std::to_string(q + 1) + " has an empty path";
        mpi = std::max(mpi, (int)npi[q]);
    }
    M.mpi = mpi;
    const int nis = std::max(isles, 1);
    M.npi.assign(nis, 0);
    M.lpisl.assign((size_t)mpi * nis, 0);
    M.ipisl.assign((size_t)mpi * nis, 0);
    M.jpisl.assign((size_t)mpi * nis, 0);
    size_t o = 0;
    for (int q = 0; q < isles; q++) {
        M.npi[q] = npi[q];
        for (int p = 0; p < npi[q]; p++, o++) {
            const int lpi = paths[3 * o], ipi = paths[3 * o + 1], jpi = paths[3 * o + 2];
            if ((std::abs(lpi) != 1 && std::abs(lpi) != 2) || ipi > mi || ipi < 1 || jpi > mj || jpi < 1 ||
                M.K1(ipi, jpi) > mk || (std::abs(lpi) == 2 && jpi + 1 > mj))
                return "island " + std::to_string(q + 1) + ": bad path point " + std::to_string(p + 1);
            M.lpisl[p + (size_t)mpi * q] = lpi;
            M.ipisl[p + (size_t)mpi * q] = ipi;
            M.jpisl[p + (size_t)mpi * q] = jpi;
        }
    }
    // the factorisation and the island basis (1788-1814): the velocity field
    // of psi = 1 on island isol and 0 elsewhere, and the path integrals of
    // each basis field round every island
    mom_band(M);
    const size_t nub = (size_t)2 * (mi + 2) * (mj + 1), nps = (size_t)(mi + 1) * (mj + 1);
    M.ubisl.assign(nub * nis, 0.0);
    M.psisl.assign(nps * nis, 0.0);
    M.erisl.assign((size_t)nis * (nis + 1), 0.0);
    std::vector<double> gb(M.N);
    for (int isol = 1; isol <= isles; isol++) {
        for (int j = 0; j <= mj; j++)
            for (int i = 1; i <= mi; i++) gb[(i + j * mi) - 1] = labels[(i - 1) + (size_t)mi * j] == isol + 1 ? 1.0 : 0.0;
        mom_solve(M, gb, M.ubisl.data() + nub * (isol - 1), M.psisl.data() + nps * (isol - 1));
        for (int isl = 1; isl <= isles; isl++)
            M.erisl[(isl - 1) + (size_t)isles * (isol - 1)] = mom_path_integral(M, M.ubisl.data() + nub * (isol - 1), isl);
    }
    // matinv_gold: the elimination of the isles x isles system, in place
    auto A = [&](int i, int j) -> double& { return M.erisl[(i - 1) + (size_t)isles * (j - 1)]; };
    for (int i = 1; i <= isles - 1; i++)
        for (int j = i + 1; j <= isles; j++)
            for (int k = i + 1; k <= isles; k++) A(j, k) = A(i, i) * A(j, k) - A(j, i) * A(i, k);
    return "";
}

}  // namespace cgo
