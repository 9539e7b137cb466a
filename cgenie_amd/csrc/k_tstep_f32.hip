// k_tstep_f32.hip -- FP32 instantiation of the flux kernel (the FP32 form:
// FP32 face fluxes, the update accumulated with its rounding residual carried).
#include "tstep_impl.h"

namespace cgo {

void launch_flux32(const MemberDev* t, const int32_t* l, int n, int maxcol, int maxL, int maxi, int lb, hipStream_t s) {
    const int tw = FLUX_THREADS + 2 * maxi;
    // keep the workgroup within 64 KiB of LDS (wide grids take smaller tracer blocks)
    while (lb > 2 && flux_lds_bytes(tw, lb, sizeof(float)) > 65536) lb = lb == 8 ? 6 : (lb == 6 ? 4 : 2);
    const dim3 grid((maxcol + FLUX_THREADS - 1) / FLUX_THREADS, (maxL + lb - 1) / lb, n);
    const size_t lds = flux_lds_bytes(tw, lb, sizeof(float));
    if (lb == 2)
        hipLaunchKernelGGL((k_flux<float, 2>), grid, dim3(FLUX_THREADS), lds, s, t, l, tw);
    else if (lb == 8)
        hipLaunchKernelGGL((k_flux<float, 8>), grid, dim3(FLUX_THREADS), lds, s, t, l, tw);
    else if (lb == 6)
        hipLaunchKernelGGL((k_flux<float, 6>), grid, dim3(FLUX_THREADS), lds, s, t, l, tw);
    else
        hipLaunchKernelGGL((k_flux<float, 4>), grid, dim3(FLUX_THREADS), lds, s, t, l, tw);
}

}  // namespace cgo
