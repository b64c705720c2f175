// sw2d_fused32.hip -- FP32 instantiation of the fused SW2D kernel (precision
// form: FP32 state, fluxes and local arithmetic, the water-surface tests on
// FP32 bed steps formed once from the FP64 bed; FP64 time, CFL minimum,
// volume ledger and hydrograph).  Built with
// -ffp-contract=on: a multiply-add is fused only within one source
// expression, so the arithmetic does not depend on the code around it (the
// kernel's instantiations give a scenario the same results).
#include "sw2d_fused_kernel.hpp"

namespace sw2d {

// The FP32 form's bed steps (built once per scenario): for every cell, the
// FP32 rounding of the FP64 bed difference to its north and to its west
// neighbour, as the kernel would form them from z; 0 toward an inactive or
// missing neighbour (no face flux reads it: both cells must be active) and
// NaN for an inactive cell, which marks it inactive.
__global__ void __launch_bounds__(256) bed_steps(const double *z, u32x2 *dz, int R, int C, int P)
{
    const long n = (long)R * P;
    for (long o = (long)blockIdx.x * blockDim.x + threadIdx.x; o < n; o += (long)gridDim.x * blockDim.x) {
        const int r = (int)(o / P), j = (int)(o % P);
        const double zc = z[o];
        float sn = __builtin_nanf(""), sw = __builtin_nanf("");
        if (j < C && zc == zc) {
            const double zn = r > 0 ? z[o - P] : __builtin_nan("");
            const double zw = j > 0 ? z[o - 1] : __builtin_nan("");
            sn = (zn == zn) ? (float)(zn - zc) : 0.0f;
            sw = (zw == zw) ? (float)(zw - zc) : 0.0f;
        }
        u32x2 v;
        v.x = __float_as_uint(sn);
        v.y = __float_as_uint(sw);
        dz[o] = v;
    }
}

hipError_t launch_fused_f32(hipStream_t stream, const Scenario *scen, const TileRef *tiles, int nblocks,
                           int phase, int features, size_t lds)
{
    return launch_fused_impl<float, float>(stream, scen, tiles, nblocks, phase, features, lds);
}
int fused_waves_per_cu_f32(int features) { return fused_waves_per_cu_impl<float, float>(features); }
hipError_t launch_init_f32(sw2d_handle *h) { return launch_init_impl<float, float>(h); }
hipError_t launch_bed_steps(sw2d_handle *h)
{
    const long n = (long)h->R * h->P;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    hipLaunchKernelGGL(bed_steps, dim3(blocks), dim3(256), 0, h->stream, (const double *)h->z,
                       (u32x2 *)h->dz, h->R, h->C, h->P);
    return hipGetLastError();
}
hipError_t launch_derive_f32(sw2d_handle *h, int field, double *out, long first, long count)
{
    return launch_derive_impl<float, float>(h, field, out, first, count);
}

}  // namespace sw2d
