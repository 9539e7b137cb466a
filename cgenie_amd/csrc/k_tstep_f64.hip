// k_tstep_f64.hip -- FP64 instantiations of the GOLDSTEIN tracer-step
// kernels, the end-of-step flip and the inventory reduction.  Built with
// -ffp-contract=off (see tstep_impl.h).
#define CGO_DEFINE_FLIP
#include "tstep_impl.h"

namespace cgo {

void launch_flux64(const MemberDev* t, const int32_t* l, int n, int maxcol, int maxL, int maxi, int lb, int nt, size_t lds_cap,
                   hipStream_t s) {
    if (nt != 256) nt = 128;
    const int tw = nt + 2 * maxi;
    // keep the workgroup within the LDS the device grants one block
    while (lb > 2 && flux_lds_bytes(tw, lb, sizeof(double), nt) > lds_cap) lb = lb == 8 ? 6 : (lb == 6 ? 4 : 2);
    if (nt == 256 && flux_lds_bytes(tw, lb, sizeof(double), nt) > lds_cap) nt = 128;
    const int tw2 = nt + 2 * maxi;
    const dim3 grid((maxcol + nt - 1) / nt, (maxL + lb - 1) / lb, n);
    const size_t lds = flux_lds_bytes(tw2, lb, sizeof(double), nt);
#define CGO_FLUX(LB_, NT_) hipLaunchKernelGGL((k_flux<double, LB_, NT_>), grid, dim3(NT_), lds, s, t, l, tw2)
    if (nt == 256) {
        if (lb == 2) CGO_FLUX(2, 256); else if (lb == 8) CGO_FLUX(8, 256); else if (lb == 6) CGO_FLUX(6, 256); else CGO_FLUX(4, 256);
    } else {
        if (lb == 2) CGO_FLUX(2, 128); else if (lb == 8) CGO_FLUX(8, 128); else if (lb == 6) CGO_FLUX(6, 128); else CGO_FLUX(4, 128);
    }
#undef CGO_FLUX
}

void launch_co(const MemberDev* t, const int32_t* l, int n, int nwetmax, int prec, hipStream_t s) {
    const dim3 grid((nwetmax + 63) / 64, n);
    if (prec == 0)
        hipLaunchKernelGGL(k_co<double>, grid, dim3(64), 0, s, t, l);
    else
        hipLaunchKernelGGL(k_co<float>, grid, dim3(64), 0, s, t, l);
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

void launch_inventory(const MemberDev* t, int32_t m, double* out, const double* w, int prec, hipStream_t s) {
    // one block per tracer; the member's L bounds the grid via the caller's buffer
    MemberDev h;
    hipMemcpy(&h, t + m, sizeof(MemberDev), hipMemcpyDeviceToHost);
    if (prec == 0)
        hipLaunchKernelGGL(k_inventory<double>, dim3(h.L), dim3(256), 0, s, t, m, out, w);
    else
        hipLaunchKernelGGL(k_inventory<float>, dim3(h.L), dim3(256), 0, s, t, m, out, w);
}

}  // namespace cgo
