// k_tstep_f64.hip -- FP64 instantiations of the GOLDSTEIN tracer-step
// kernels, the end-of-step flip and the inventory reduction.  Built with
// -ffp-contract=off (see tstep_impl.h).
#define CGO_DEFINE_FLIP
#include "tstep_impl.h"

namespace cgo {

void launch_flux64(const MemberDev* t, const int32_t* l, int n, const FluxShape& fs, int maxL, int lb, int nt,
                   int nbuf, int iso_split, int part, size_t lds_cap, hipStream_t s) {
    launch_flux_t<double>(t, l, n, fs, maxL, lb, nt, nbuf, iso_split, part, lds_cap, s);
}

void launch_co(const MemberDev* t, const int32_t* l, int n, int nwetmax, int maxkmax, int prec, hipStream_t s) {
    // 64 wet columns per workgroup of 128 threads (32 convecting at a time,
    // a group taking a second column when it has finished its first): small
    // workgroups, many resident per CU whether few columns convect or most;
    // fewer threads where the level arrays would not fit in 64 KiB of LDS
    const int K = maxkmax + 2;
    int nt = 128;
    while (nt > 32 && co_lds_bytes(nt, nt / 2, K) > 65536) nt /= 2;
    const int cb = nt / 2;
    const dim3 grid((nwetmax + cb - 1) / cb, n);
    const size_t lds = co_lds_bytes(nt, cb, K);
    if (prec == 0)
        hipLaunchKernelGGL(k_co<double>, grid, dim3(nt), lds, s, t, l, K, cb);
    else
        hipLaunchKernelGGL(k_co<float>, grid, dim3(nt), lds, s, t, l, K, cb);
}

void launch_flip(const MemberDev* t, const int32_t* l, int n, hipStream_t s) {
    hipLaunchKernelGGL(k_flip, dim3((n + 63) / 64), dim3(64), 0, s, t, l, n);
}

// per-tracer volume-weighted sum over the member's current state
template <typename R>
__global__ __launch_bounds__(256) void k_inventory(const MemberDev* __restrict__ mems, int32_t m, double* out,
                                                   const double* __restrict__ w) {
    const MemberDev& M = mems[m];
    const int l = blockIdx.x;
    const int par = *gp(M.parity);
    const size_t ncell = M.g.ncell;
    const size_t base = (size_t)l * ncell;
    double s = 0.0;
    for (size_t c = threadIdx.x; c < ncell; c += 256) {
        const double wc = w[c];
        if (wc != 0.0) s += StateIO<R>::ld(M.st[par], M.lo[par], base + c) * wc;
    }
    __shared__ double red[256];
    red[threadIdx.x] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (threadIdx.x < h) red[threadIdx.x] += red[threadIdx.x + h];
        __syncthreads();
    }
    if (threadIdx.x == 0) out[l] = red[0];
}

void launch_inventory(const MemberDev* t, int32_t m, int L, double* out, const double* w, int prec, hipStream_t s) {
    // one block per tracer
    if (prec == 0)
        hipLaunchKernelGGL(k_inventory<double>, dim3(L), dim3(256), 0, s, t, m, out, w);
    else
        hipLaunchKernelGGL(k_inventory<float>, dim3(L), dim3(256), 0, s, t, m, out, w);
}

}  // namespace cgo
