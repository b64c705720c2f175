// sw2d_fused64.hip -- FP64 instantiation of the fused SW2D kernel (parity
// mode).  Built with -ffp-contract=off: no multiply-add is fused, so each
// expression rounds in the order it is written, as in the CPU FP64 check.
#include "sw2d_fused_kernel.hpp"

namespace sw2d {

__global__ void __launch_bounds__(64) peek_kernel(const StepCommon s, Ledger *out)
{
    ledger_peek(s, out, threadIdx.x);
}

hipError_t launch_fused_f64(hipStream_t stream, const Scenario *scen, const TileRef *tiles, int nblocks,
                           int phase, int features, size_t lds)
{
    return launch_fused_impl<double, double>(stream, scen, tiles, nblocks, phase, features, lds);
}
int fused_waves_per_cu_f64(int features) { return fused_waves_per_cu_impl<double, double>(features); }
hipError_t launch_init_f64(sw2d_handle *h) { return launch_init_impl<double, double>(h); }
hipError_t launch_derive_f64(sw2d_handle *h, int field, double *out, long first, long count)
{
    return launch_derive_impl<double, double>(h, field, out, first, count);
}

hipError_t launch_peek(sw2d_handle *h, const StepCommon &sc)
{
    hipLaunchKernelGGL(peek_kernel, dim3(1), dim3(64), 0, h->stream, sc, h->peek);
    return hipGetLastError();
}

}  // namespace sw2d
