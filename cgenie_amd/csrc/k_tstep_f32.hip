// k_tstep_f32.hip -- FP32 instantiation of the flux kernel (the FP32 form:
// FP32 face fluxes, the update accumulated with its rounding residual carried).
#include "tstep_impl.h"

namespace cgo {

void launch_flux32(const MemberDev* t, const int32_t* l, int n, int maxcol, int maxL, int maxi, int lb, int nt, size_t lds_cap,
                   hipStream_t s) {
    if (nt != 256) nt = 128;
    const int tw = nt + 2 * maxi;
    // keep the workgroup within the LDS the device grants one block
    while (lb > 2 && flux_lds_bytes(tw, lb, sizeof(float), nt) > lds_cap) lb = lb == 8 ? 6 : (lb == 6 ? 4 : 2);
    if (nt == 256 && flux_lds_bytes(tw, lb, sizeof(float), nt) > lds_cap) nt = 128;
    const int tw2 = nt + 2 * maxi;
    const dim3 grid((maxcol + nt - 1) / nt, (maxL + lb - 1) / lb, n);
    const size_t lds = flux_lds_bytes(tw2, lb, sizeof(float), nt);
#define CGO_FLUX(LB_, NT_) hipLaunchKernelGGL((k_flux<float, LB_, NT_>), grid, dim3(NT_), lds, s, t, l, tw2)
    if (nt == 256) {
        if (lb == 2) CGO_FLUX(2, 256); else if (lb == 8) CGO_FLUX(8, 256); else if (lb == 6) CGO_FLUX(6, 256); else CGO_FLUX(4, 256);
    } else {
        if (lb == 2) CGO_FLUX(2, 128); else if (lb == 8) CGO_FLUX(8, 128); else if (lb == 6) CGO_FLUX(6, 128); else CGO_FLUX(4, 128);
    }
#undef CGO_FLUX
}

}  // namespace cgo
