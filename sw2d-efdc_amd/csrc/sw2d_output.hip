// sw2d_output.hip -- asynchronous output copies of a stand-alone handle
// (sw2d_snapshot* in include/sw2d.h).
//
// The reference stops at every output interval to copy h, uu, vv (and the
// coupling velocities) to the host with blocking copies
// (SW2D-EFDC.cu:4880-4916) and writes the files before stepping on.  Here an
// output is a snapshot: the fields are derived into device buffers on the
// compute stream (stream order puts them after the steps before and before
// the steps after), then copied into pinned host buffers by a copy stream
// that waits for them, so the DMA overlaps the following steps.  Two slots
// give double buffering: a writer formats one snapshot while the next is in
// flight; a third snapshot waits until a slot is released.
#include <condition_variable>
#include <memory>
#include <mutex>

#include "sw2d_ops.hpp"

using namespace sw2d;

namespace sw2d {

constexpr int kSnapFields = SW2D_FIELD_VV1 + 1;
constexpr int kSnapSlots = 2;

struct SnapSlot {
    long ticket = -1;                      // -1: free
    unsigned mask = 0;
    bool nodata_done = false;              // z: NaN marks replaced by the grid's nodata
    double *dev[kSnapFields] = {};
    double *host[kSnapFields] = {};        // pinned
    hipEvent_t ready = nullptr;            // fields derived (compute stream)
    hipEvent_t copied = nullptr;           // fields on the host (copy stream)
};

struct SnapRing {
    std::mutex mu;
    std::condition_variable freed;
    hipStream_t copy = nullptr;
    SnapSlot slot[kSnapSlots];
    long next = 0;
};

void snapshot_free(sw2d_handle *h)
{
    SnapRing *r = h->snap;
    if (!r) return;
    if (r->copy) (void)hipStreamSynchronize(r->copy);
    for (SnapSlot &s : r->slot) {
        for (int f = 0; f < kSnapFields; f++) {
            if (s.dev[f]) (void)hipFree(s.dev[f]);
            if (s.host[f]) (void)hipHostFree(s.host[f]);
        }
        if (s.ready) (void)hipEventDestroy(s.ready);
        if (s.copied) (void)hipEventDestroy(s.copied);
    }
    if (r->copy) (void)hipStreamDestroy(r->copy);
    delete r;
    h->snap = nullptr;
}

}  // namespace sw2d

namespace {

// plain-path device array of a field (FP64, R*C from the flat layout's start)
const double *plain_field(const sw2d_handle *h, int field)
{
    switch (field) {
    case SW2D_FIELD_H: return h->ph;
    case SW2D_FIELD_UM: return h->pum;
    case SW2D_FIELD_VN: return h->pvn;
    case SW2D_FIELD_UU: return h->puu;
    case SW2D_FIELD_VV: return h->pvv;
    case SW2D_FIELD_UUA: return h->puua;
    case SW2D_FIELD_VVA: return h->pvva;
    case SW2D_FIELD_Z: return h->pz;
    case SW2D_FIELD_HM: return h->phm;
    case SW2D_FIELD_HN: return h->phn;
    case SW2D_FIELD_UU1: return h->puu1;
    case SW2D_FIELD_VV1: return h->pvv1;
    }
    return nullptr;
}

SnapSlot *find_slot(SnapRing *r, long ticket)
{
    for (SnapSlot &s : r->slot)
        if (s.ticket == ticket && ticket >= 0) return &s;
    return nullptr;
}

}  // namespace

extern "C" {

int sw2d_snapshot(sw2d_handle *h, unsigned fields, double timeout_s, long *ticket)
{
    if (!h || !ticket) return SW2D_E_ARG;
    if (fields == 0 || (fields >> kSnapFields) != 0) return fail(h, SW2D_E_ARG, "bad field mask");
    if (h->batch) return fail(h, SW2D_E_STATE, "snapshots are for stand-alone handles");
    if (!h->snap) {
        SnapRing *r = new SnapRing();
        hipError_t e = hipStreamCreate(&r->copy);
        for (SnapSlot &s : r->slot) {
            if (e == hipSuccess) e = hipEventCreateWithFlags(&s.ready, hipEventDisableTiming);
            if (e == hipSuccess) e = hipEventCreateWithFlags(&s.copied, hipEventDisableTiming);
        }
        h->snap = r;
        if (e != hipSuccess) {
            snapshot_free(h);
            return hip_fail(h, e, "snapshot set-up");
        }
    }
    SnapRing *r = h->snap;
    SnapSlot *s = nullptr;
    {
        std::unique_lock<std::mutex> lk(r->mu);
        auto have = [&] {
            for (SnapSlot &x : r->slot)
                if (x.ticket < 0) { s = &x; return true; }
            return false;
        };
        const auto limit = std::chrono::duration<double>(timeout_s > 0 ? timeout_s : 0.0);
        if (!r->freed.wait_for(lk, limit, have))
            return fail(h, SW2D_E_STATE, "both snapshot slots are held: release one first");
        s->ticket = r->next++;   // taken: only this thread fills it
    }
    const size_t RC = (size_t)h->R * h->C;
    auto give_back = [&](int code) {
        std::lock_guard<std::mutex> lk(r->mu);
        s->ticket = -1;
        r->freed.notify_all();
        return code;
    };
    for (int f = 0; f < kSnapFields; f++) {
        if (!(fields >> f & 1u)) continue;
        hipError_t e = hipSuccess;
        if (!s->dev[f]) e = hipMalloc(&s->dev[f], RC * 8);
        if (e == hipSuccess && !s->host[f]) e = hipHostMalloc(&s->host[f], RC * 8, 0);
        if (e != hipSuccess) return give_back(fail(h, SW2D_E_NOMEM, std::string("snapshot buffers: ") + hipGetErrorString(e)));
    }
    // the copy stream may still be reading this slot's device buffers for
    // the previous snapshot in it: the released ticket has waited for that
    hipError_t e = hipSuccess;
    for (int f = 0; f < kSnapFields && e == hipSuccess; f++) {
        if (!(fields >> f & 1u)) continue;
        if (fused(h))
            e = launch_derive(h, f, s->dev[f], 0, (long)RC);
        else
            e = hipMemcpyAsync(s->dev[f], plain_field(h, f), RC * 8, hipMemcpyDeviceToDevice, h->stream);
    }
    if (e == hipSuccess) e = hipEventRecord(s->ready, h->stream);
    if (e == hipSuccess) e = hipStreamWaitEvent(r->copy, s->ready, 0);
    for (int f = 0; f < kSnapFields && e == hipSuccess; f++)
        if (fields >> f & 1u) e = hipMemcpyAsync(s->host[f], s->dev[f], RC * 8, hipMemcpyDeviceToHost, r->copy);
    if (e == hipSuccess) e = hipEventRecord(s->copied, r->copy);
    if (e != hipSuccess) {
        (void)hipStreamSynchronize(r->copy);
        return give_back(hip_fail(h, e, "snapshot"));
    }
    s->mask = fields;
    s->nodata_done = false;
    *ticket = s->ticket;
    return SW2D_OK;
}

int sw2d_snapshot_field(sw2d_handle *h, long ticket, int field, const double **data)
{
    if (!h || !data) return SW2D_E_ARG;
    *data = nullptr;
    SnapRing *r = h->snap;
    if (!r) return SW2D_E_STATE;
    SnapSlot *s;
    {
        std::lock_guard<std::mutex> lk(r->mu);
        s = find_slot(r, ticket);
    }
    if (!s) return SW2D_E_STATE;
    if (field < 0 || field >= kSnapFields || !(s->mask >> field & 1u)) return SW2D_E_ARG;
    if (hipEventSynchronize(s->copied) != hipSuccess) return SW2D_E_HIP;
    if (field == SW2D_FIELD_Z && fused(h)) {
        std::lock_guard<std::mutex> lk(r->mu);
        if (!s->nodata_done) {
            const size_t RC = (size_t)h->R * h->C;
            double *z = s->host[field];
            for (size_t i = 0; i < RC; i++)
                if (z[i] != z[i]) z[i] = h->grid.nodata;
            s->nodata_done = true;
        }
    }
    *data = s->host[field];
    return SW2D_OK;
}

int sw2d_snapshot_release(sw2d_handle *h, long ticket)
{
    if (!h) return SW2D_E_ARG;
    SnapRing *r = h->snap;
    if (!r) return SW2D_E_STATE;
    SnapSlot *s;
    {
        std::lock_guard<std::mutex> lk(r->mu);
        s = find_slot(r, ticket);
    }
    if (!s) return SW2D_E_STATE;
    // the slot's buffers are reused only after its copies have finished
    const hipError_t e = hipEventSynchronize(s->copied);
    std::lock_guard<std::mutex> lk(r->mu);
    s->ticket = -1;
    r->freed.notify_all();
    return e == hipSuccess ? SW2D_OK : SW2D_E_HIP;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// sw2d_get_field_noalloc: sw2d_get_field staged through a fixed device array
// instead of a buffer the library allocates.  The array is a file-scope
// device variable: every device the library runs on holds its own copy, at
// its own address, from module load until the process ends.  A field is
// derived into it and copied out in row chunks, so any grid whose row fits
// is served.  Host code never passes the variable's name to a kernel or a
// copy (in host code it names a host-side stand-in): it takes the device
// address from hipGetSymbolAddress, once per device, with that device
// current.
// ---------------------------------------------------------------------------

namespace {

constexpr size_t kNoAllocDoubles = (size_t)2 << 20;   // 16 MiB of FP64 values per device

static __device__ __align__(256) double scratch[kNoAllocDoubles];

struct NoAllocDevice {
    double *ptr = nullptr;   // device address of `scratch` on this device
    std::mutex use;          // one call at a time on one device
};

std::mutex g_noalloc_mu;
std::vector<std::unique_ptr<NoAllocDevice>> g_noalloc;

// the device's staging area; looked up on the first call on that device
// (the caller has made `device` current)
int noalloc_area(sw2d_handle *h, int device, NoAllocDevice **out)
{
    std::lock_guard<std::mutex> lk(g_noalloc_mu);
    if ((int)g_noalloc.size() <= device) g_noalloc.resize(device + 1);
    if (!g_noalloc[device]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(h, e, "hipGetSymbolAddress(scratch)");
        g_noalloc[device] = std::make_unique<NoAllocDevice>();
        g_noalloc[device]->ptr = (double *)p;
    }
    *out = g_noalloc[device].get();
    return SW2D_OK;
}

}  // namespace

extern "C" {

int sw2d_get_field_noalloc(sw2d_handle *h, int field, double *out)
{
    if (!h) return SW2D_E_ARG;
    if (!out) return fail(h, SW2D_E_ARG, "out is NULL");
    if (field < 0 || field > SW2D_FIELD_VV1) return fail(h, SW2D_E_ARG, "unknown field");
    if (fused(h) && (size_t)h->C > kNoAllocDoubles)
        return fail(h, SW2D_E_ARG, "grid rows wider than the 2 Mi-value staging area of "
                                   "sw2d_get_field_noalloc; use sw2d_get_field");
    int prev = 0;
    HIPCHK(h, hipGetDevice(&prev));
    HIPCHK(h, hipSetDevice(h->device));
    int rc = SW2D_OK;
    const size_t RC = (size_t)h->R * h->C;
    hipError_t e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) {
        rc = hip_fail(h, e, "hipStreamSynchronize");
    } else if (!fused(h)) {
        // the plain layout holds every field in a device array already
        e = hipMemcpy(out, plain_field(h, field), RC * 8, hipMemcpyDeviceToHost);
        if (e != hipSuccess) rc = hip_fail(h, e, "hipMemcpy");
    } else {
        NoAllocDevice *area = nullptr;
        rc = noalloc_area(h, h->device, &area);
        if (rc == SW2D_OK) {
            std::lock_guard<std::mutex> lk(area->use);
            // whole rows per chunk; chunk k + 1 is derived after the copy of
            // chunk k in stream order, so one staging area serves them all
            const long rows_per = (long)(kNoAllocDoubles / (size_t)h->C);
            for (long r0 = 0; r0 < h->R && e == hipSuccess; r0 += rows_per) {
                const long nr = std::min<long>(rows_per, h->R - r0);
                const long first = r0 * h->C, count = nr * h->C;
                e = launch_derive(h, field, area->ptr, first, count);
                if (e == hipSuccess)
                    e = hipMemcpyAsync(out + first, area->ptr, (size_t)count * 8,
                                       hipMemcpyDeviceToHost, h->stream);
            }
            const hipError_t e2 = hipStreamSynchronize(h->stream);
            if (e == hipSuccess) e = e2;
            if (e != hipSuccess) rc = hip_fail(h, e, "get_field_noalloc");
        }
        if (rc == SW2D_OK && field == SW2D_FIELD_Z)
            for (size_t i = 0; i < RC; i++)
                if (out[i] != out[i]) out[i] = h->grid.nodata;
    }
    (void)hipSetDevice(prev);
    return rc;
}

}  // extern "C"
