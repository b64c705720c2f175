// sw2d_fused32.hip -- FP32 instantiation of the fused SW2D kernel (precision
// form: FP32 state, fluxes and local arithmetic; FP64 bed elevation, water
// surfaces, time, CFL minimum, volume ledger and hydrograph).  Built with
// -ffp-contract=on: a multiply-add is fused only within one source
// expression, so the arithmetic does not depend on the code around it (the
// kernel's instantiations give a scenario the same results).
#include "sw2d_fused_kernel.hpp"

namespace sw2d {

hipError_t launch_fused_f32(hipStream_t stream, const Scenario *scen, const TileRef *tiles, int nblocks,
                           int phase, int features, size_t lds)
{
    return launch_fused_impl<float, float>(stream, scen, tiles, nblocks, phase, features, lds);
}
int fused_waves_per_cu_f32(int features) { return fused_waves_per_cu_impl<float, float>(features); }
hipError_t launch_init_f32(sw2d_handle *h) { return launch_init_impl<float, float>(h); }
hipError_t launch_derive_f32(sw2d_handle *h, int field, double *out, long first, long count)
{
    return launch_derive_impl<float, float>(h, field, out, first, count);
}

}  // namespace sw2d
