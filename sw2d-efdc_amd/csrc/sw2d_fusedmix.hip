// sw2d_fusedmix.hip -- mixed instantiation of the fused SW2D kernel: FP32
// fluxes and momentum arithmetic (as in the FP32 form) over FP64 depths.  The
// depths, both continuity updates, the face depths' dry tests, the weir
// switches (water surface against the neighbour's bed) and treat_error's dry
// test are FP64, so every depth-driven switch sees the depths the FP64 form
// sees; only the fluxes carry FP32 rounding.  Built with -ffp-contract=on, as
// the FP32 form.
#include "sw2d_fused_kernel.hpp"

namespace sw2d {

hipError_t launch_fused_mix(hipStream_t stream, const Scenario *scen, const TileRef *tiles, int nblocks,
                            int phase, int features, size_t lds)
{
    return launch_fused_impl<float, double>(stream, scen, tiles, nblocks, phase, features, lds);
}
int fused_waves_per_cu_mix(int features) { return fused_waves_per_cu_impl<float, double>(features); }
hipError_t launch_init_mix(sw2d_handle *h) { return launch_init_impl<float, double>(h); }
hipError_t launch_derive_mix(sw2d_handle *h, int field, double *out, long first, long count)
{
    return launch_derive_impl<float, double>(h, field, out, first, count);
}

}  // namespace sw2d
