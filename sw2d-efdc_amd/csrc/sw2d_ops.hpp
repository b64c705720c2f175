// sw2d_ops.hpp -- host-side helpers shared by the handle (sw2d_api.hip) and
// batch (sw2d_batch.hip) implementations of the C ABI.
#pragma once
#include <math.h>
#include <string.h>

#include <algorithm>
#include <functional>
#include <new>
#include <string>
#include <vector>

#include "sw2d_internal.hpp"

namespace sw2d {

inline thread_local std::string g_create_error;

inline int fail(sw2d_handle *h, int code, const std::string &msg)
{
    if (h) h->err = msg;
    else g_create_error = msg;
    return code;
}

inline int hip_fail(sw2d_handle *h, hipError_t e, const char *what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(h, SW2D_E_HIP, m);
}

#define HIPCHK(h, expr)                                   \
    do {                                                  \
        hipError_t e_ = (expr);                           \
        if (e_ != hipSuccess) return hip_fail(h, e_, #expr); \
    } while (0)

// device arrays of a handle: hipMalloc, or the pool of its batch; zeroed
// in stream order (a pool range freed by a retired scenario is reused only
// after the launches queued before it)
template <typename T>
inline int dalloc(sw2d_handle *h, T **p, size_t count)
{
    void *q = nullptr;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e;
    if (h->batch) {
        q = h->batch->pool.alloc(bytes);
        if (!q) return fail(h, SW2D_E_NOMEM, "batch memory pool exhausted");
    } else {
        e = hipMalloc(&q, bytes);
        if (e != hipSuccess) return fail(h, SW2D_E_NOMEM, std::string("hipMalloc: ") + hipGetErrorString(e));
    }
    e = hipMemsetAsync(q, 0, bytes, h->stream);
    if (e != hipSuccess) return hip_fail(h, e, "hipMemsetAsync");
    h->allocs.push_back(q);
    *p = (T *)q;
    return SW2D_OK;
}

inline void dfree(sw2d_handle *h, void *p)
{
    if (!p) return;
    auto it = std::find(h->allocs.begin(), h->allocs.end(), p);
    if (it != h->allocs.end()) h->allocs.erase(it);
    if (h->batch) (void)h->batch->pool.release(p);
    else (void)hipFree(p);
}

inline void drop_graph(sw2d_handle *h)
{
    if (h->gexec) (void)hipGraphExecDestroy(h->gexec);
    h->gexec = nullptr;
}

// a setter changed what the kernels read: refresh the descriptor before the
// next launch and capture the graph again (its grid size may have changed)
inline void changed(sw2d_handle *h)
{
    h->desc_dirty = true;
    drop_graph(h);
    if (h->batch) h->batch->dirty = true;
}

// the fused kernel's initial CFL minimum and derived fields, in the
// handle's precision form
inline hipError_t launch_init(sw2d_handle *h)
{
    const int p = h->prm.precision;
    return p == SW2D_FP64 ? launch_init_f64(h) : p == SW2D_FP32 ? launch_init_f32(h) : launch_init_mix(h);
}
inline hipError_t launch_derive(sw2d_handle *h, int field, double *out_dev, long first, long count)
{
    const int p = h->prm.precision;
    if (p == SW2D_FP64) return launch_derive_f64(h, field, out_dev, first, count);
    if (p == SW2D_FP32) return launch_derive_f32(h, field, out_dev, first, count);
    return launch_derive_mix(h, field, out_dev, first, count);
}

// defined in sw2d_api.hip
int round_up(int x, int m);
StepCommon make_common(sw2d_handle *h);
bool fused(const sw2d_handle *h);
int fused_blocks(const sw2d_handle *h);
Scenario make_scenario(sw2d_handle *h);
int choose_tile_rows(const sw2d_handle *h);
long wave_slots(int device, int prec, int features);
int blocks_at(const sw2d_handle *h, int H);
std::vector<int> round_fill_heights(long slots, const std::function<long(int)> &blocks);
void set_tiles(sw2d_handle *h, int H);
int retile(sw2d_handle *h, int H);
void advance(sw2d_handle *h, long m);
// Host images of a fused-kernel scenario's inputs in the layout of its
// device arrays: the bed in FP64 with NaN for inactive cells and the initial
// depths in the form's depth type, both at pitch P (columns rounded up to
// 16), plus the initial volume.  create_impl builds them itself unless it is
// given them: a server builds them on the submitting thread, so that its
// device thread, which keeps the GPU fed, only copies them.
struct HostImages {
    int P = 0;
    double volume0 = 0;           // sum of h0 over active cells (row-major order), x dx^2
    std::vector<double> z;        // R x P
    std::vector<double> h64;      // R x P, FP64 depth forms
    std::vector<float> h32;       // R x P, FP32 form
};
void prepare_images(const sw2d_grid_desc *grid, const double *z, const double *h0, int precision,
                    HostImages &img);
int create_impl(const sw2d_grid_desc *grid, const double *z, const double *h0,
                const sw2d_params *params, sw2d_batch *batch, sw2d_handle **out,
                const HostImages *img = nullptr);
// defined in sw2d_output.hip
void snapshot_free(sw2d_handle *h);
// defined in sw2d_batch.hip
int batch_fit_tiles(sw2d_batch *b, int *tile_rows, int *retiled);

}  // namespace sw2d
