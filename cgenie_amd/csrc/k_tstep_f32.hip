// k_tstep_f32.hip -- FP32 instantiation of the flux kernel (the FP32 form:
// FP32 face fluxes, the update accumulated with its rounding residual carried).
#include "tstep_impl.h"

namespace cgo {

void launch_flux32(const MemberDev* t, const int32_t* l, int n, const FluxShape& fs, int maxL, int lb, int nt,
                   int nbuf, int iso_split, int part, size_t lds_cap, hipStream_t s) {
    launch_flux_t<float>(t, l, n, fs, maxL, lb, nt, nbuf, iso_split, part, lds_cap, s);
}

}  // namespace cgo
