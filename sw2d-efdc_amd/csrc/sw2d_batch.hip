// sw2d_batch.hip -- batches: many scenarios stepped by one launch per
// iteration (C ABI in include/sw2d.h).
//
// Each member is an ordinary handle whose arrays come from the batch's memory
// pool and whose work runs on the batch's stream.  A launch covers the blocks
// of every attached member: block b of the grid reads entry b of a tile table
// (scenario slot, block within the scenario) and the scenario's descriptor,
// so members of different sizes, forcing and time steps share one grid.
// Members join at any launch: a member's iteration k runs at launch phase
// (phase0 + k) mod 6, and each member stops by itself (a done flag its own
// bookkeeping wave sets) when its iteration or time bound is reached.
#include <math.h>
#include <stdio.h>

#include "sw2d_common.hpp"
#include "sw2d_ops.hpp"

using namespace sw2d;

namespace {

int bfail(sw2d_batch *b, int code, const std::string &msg)
{
    if (b) b->err = msg;
    else g_create_error = msg;
    return code;
}

int bhip(sw2d_batch *b, hipError_t e, const char *what)
{
    return bfail(b, SW2D_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

#define BCHK(b, expr)                                   \
    do {                                                \
        hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return bhip(b, e_, #expr); \
    } while (0)

void bdrop_graph(sw2d_batch *b)
{
    if (b->gexec) (void)hipGraphExecDestroy(b->gexec);
    b->gexec = nullptr;
}

hipError_t blaunch(sw2d_batch *b, const Scenario *scen, const TileRef *tiles, int nblocks, int phase,
                   int features)
{
    return launch_fused(b->precision, b->stream, scen, tiles, nblocks, phase, features, b->lds_pad);
}

// unused LDS per one-wave block that leaves room for at most `cap` resident
// blocks per SIMD (four SIMDs per CU); 0 for no cap
size_t lds_for_cap(int device, int cap)
{
    if (cap <= 0) return 0;
    hipDeviceProp_t prop;
    if (hipGetDeviceProperties(&prop, device) != hipSuccess) return 0;
    size_t per = prop.maxSharedMemoryPerMultiProcessor / (4 * (size_t)cap);
    per = per / 256 * 256;
    return std::min<size_t>(per, prop.sharedMemPerBlock);
}

// the launch's tile table: each member's blocks in a run (order 0), or the
// members interleaved block by block (order 1), so that every stretch of
// the launch holds tiles of many members
void order_tiles(const std::vector<int> &counts, const std::vector<int> &sids, int order,
                 std::vector<TileRef> &tiles)
{
    tiles.clear();
    if (order == 0) {
        for (size_t m = 0; m < counts.size(); m++)
            for (int i = 0; i < counts[m]; i++) tiles.push_back(TileRef{sids[m], i});
        return;
    }
    int most = 0;
    for (int c : counts) most = std::max(most, c);
    for (int i = 0; i < most; i++)
        for (size_t m = 0; m < counts.size(); m++)
            if (i < counts[m]) tiles.push_back(TileRef{sids[m], i});
}

// descriptors and tile table of the attached members
int rebuild(sw2d_batch *b)
{
    if (!b->dirty) return SW2D_OK;
    BCHK(b, hipStreamSynchronize(b->stream));
    std::vector<Scenario> desc;
    std::vector<TileRef> tiles;
    int top = 0;
    for (sw2d_handle *h : b->members) top = std::max(top, h->member_slot + 1);
    desc.resize(top);
    memset(desc.data(), 0, sizeof(Scenario) * desc.size());
    int features = 0;
    std::vector<int> counts, sids;
    for (sw2d_handle *h : b->members) {
        Scenario d = make_scenario(h);
        features |= features_of(d.s);
        d.phase0 = h->phase0;
        d.k_end = h->k_end;
        d.t_end = h->t_end;
        d.status = b->dstatus + h->member_slot;
        desc[h->member_slot] = d;
        counts.push_back(d.nblocks);
        sids.push_back(h->member_slot);
        h->desc_dirty = true;  // its own descriptor is refreshed when it steps alone again
    }
    order_tiles(counts, sids, b->order, tiles);
    if ((int)tiles.size() > b->tiles_cap) {
        bdrop_graph(b);
        if (b->dtiles) (void)hipFree(b->dtiles);
        b->dtiles = nullptr;
        b->tiles_cap = 0;
        const int cap = std::max<int>(1024, (int)tiles.size() * 2);
        BCHK(b, hipMalloc(&b->dtiles, sizeof(TileRef) * cap));
        b->tiles_cap = cap;
    }
    if (top > 0)
        BCHK(b, hipMemcpy(b->dscen, desc.data(), sizeof(Scenario) * top, hipMemcpyHostToDevice));
    if (!tiles.empty())
        BCHK(b, hipMemcpy(b->dtiles, tiles.data(), sizeof(TileRef) * tiles.size(), hipMemcpyHostToDevice));
    if ((int)tiles.size() != b->nblocks || features != b->features) bdrop_graph(b);
    b->nblocks = (int)tiles.size();
    b->features = features;
    b->dirty = false;
    return SW2D_OK;
}

int ensure_bgraph(sw2d_batch *b)
{
    if (b->gexec) return SW2D_OK;
    hipGraph_t g = nullptr;
    BCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int ph = 0; ph < 6 && e == hipSuccess; ph++) e = blaunch(b, b->dscen, b->dtiles, b->nblocks, ph, b->features);
    hipError_t e2 = hipStreamEndCapture(b->stream, &g);
    if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return bhip(b, e, "graph capture launch"); }
    if (e2 != hipSuccess) return bhip(b, e2, "hipStreamEndCapture");
    e = hipGraphInstantiate(&b->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { b->gexec = nullptr; return bhip(b, e, "hipGraphInstantiate"); }
    return SW2D_OK;
}

// members' host mirrors from the device status records
int refresh(sw2d_batch *b, int *running)
{
    int top = 0;
    for (sw2d_handle *h : b->members) top = std::max(top, h->member_slot + 1);
    if (top > 0) {
        BCHK(b, hipMemcpyAsync(b->hstatus, b->dstatus, sizeof(Status) * top, hipMemcpyDeviceToHost, b->stream));
    }
    BCHK(b, hipStreamSynchronize(b->stream));
    int run = 0;
    for (sw2d_handle *h : b->members) {
        const Status &s = b->hstatus[h->member_slot];
        const long long dk = s.k - h->step_index;
        h->step_index = s.k;
        h->host_mstep += 2 * dk;
        h->host_time = s.t;
        h->cur = (int)(h->step_index & 1);
        if (!(s.done || s.fin)) run++;
    }
    if (running) *running = run;
    return SW2D_OK;
}

int flush_bevents(sw2d_batch *b)
{
    BCHK(b, hipStreamSynchronize(b->stream));
    for (auto &ev : b->events) {
        float ms = 0;
        BCHK(b, hipEventElapsedTime(&ms, ev.first, ev.second));
        b->kernel_ms += ms;
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    b->events.clear();
    return SW2D_OK;
}

size_t pool_bytes_of(size_t bytes) { return Pool::rounded(bytes); }

}  // namespace

namespace sw2d {

// Round-filling tile height for the members still running (a member with a
// height of its own keeps it): the fewest whole rounds of wave slots whose
// height is at most 80, else 32 when even six rounds need taller tiles (on
// the serving mix two rounds at H ~ 55 beat three at H ~ 37 by 1%: fewer halo
// rows per output row).  The
// running members are retiled to it and new members start at it.  A member
// whose retile finds no pool memory keeps its height (members of one launch
// may differ in height).  *retiled counts the members moved.
int batch_fit_tiles(sw2d_batch *b, int *tile_rows, int *retiled)
{
    if (retiled) *retiled = 0;
    int rc = refresh(b, nullptr);
    if (rc) return rc;
    std::vector<sw2d_handle *> run;
    int feats = 0;
    for (sw2d_handle *h : b->members) {
        const Status &st = b->hstatus[h->member_slot];
        if (st.done || st.fin) continue;
        run.push_back(h);
        feats |= features_of(make_common(h));
    }
    if (!run.empty()) {
        const long slots = wave_slots(b->device, b->precision, feats);
        const std::vector<int> fit = round_fill_heights(slots, [&](int H) {
            long n = 0;
            for (const sw2d_handle *h : run) n += blocks_at(h, h->prm.tile_rows > 0 ? h->prm.tile_rows : H);
            return n;
        });
        int best = 32;
        for (auto it = fit.rbegin(); it != fit.rend(); ++it)
            if (*it <= 80) { best = *it; break; }
        b->H = best;
        for (sw2d_handle *h : run) {
            if (h->prm.tile_rows > 0 || h->H == best) continue;
            rc = retile(h, best);
            if (rc == SW2D_E_NOMEM) continue;
            if (rc) { b->err = h->err; return rc; }
            if (retiled) (*retiled)++;
        }
    }
    if (tile_rows) *tile_rows = b->H;
    return SW2D_OK;
}

}  // namespace sw2d

extern "C" {

int sw2d_batch_create(int device, int precision, size_t pool_bytes, int tile_rows, sw2d_batch **out)
{
    if (!out) return bfail(nullptr, SW2D_E_ARG, "out is NULL");
    *out = nullptr;
    if (!known_precision(precision)) return bfail(nullptr, SW2D_E_ARG, "unknown precision");
    if (tile_rows < 0 || tile_rows > 512) return bfail(nullptr, SW2D_E_ARG, "tile_rows must be in [0, 512]");
    if (pool_bytes < (1u << 20)) return bfail(nullptr, SW2D_E_ARG, "pool must be at least 1 MiB");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return bfail(nullptr, SW2D_E_NODEV, "no HIP device");
    if (device < 0 || device >= ndev) return bfail(nullptr, SW2D_E_NODEV, "bad device ordinal");
    sw2d_batch *b = new sw2d_batch();
    b->graph_on = sw2d::graphs_enabled();
    b->device = device;
    b->precision = precision;
    b->H = tile_rows > 0 ? tile_rows : 32;
    auto bail = [&](hipError_t e, const char *what) {
        bhip(b, e, what);
        g_create_error = b->err;
        sw2d_batch_destroy(b);
        return SW2D_E_HIP;
    };
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) return bail(e, "hipSetDevice");
    if ((e = hipStreamCreate(&b->stream)) != hipSuccess) return bail(e, "hipStreamCreate");
    e = hipMalloc(&b->pool_base, pool_bytes);
    if (e != hipSuccess) {
        g_create_error = std::string("pool hipMalloc: ") + hipGetErrorString(e);
        sw2d_batch_destroy(b);
        return SW2D_E_NOMEM;
    }
    b->pool.init(b->pool_base, pool_bytes);
    if ((e = hipMalloc(&b->dscen, sizeof(Scenario) * kMaxMembers)) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = hipMalloc(&b->dstatus, sizeof(Status) * kMaxMembers)) != hipSuccess) return bail(e, "hipMalloc");
    if ((e = hipHostMalloc(&b->hstatus, sizeof(Status) * kMaxMembers, 0)) != hipSuccess) return bail(e, "hipHostMalloc");
    for (int i = kMaxMembers - 1; i >= 0; i--) b->free_slots.push_back(i);
    *out = b;
    return SW2D_OK;
}

int sw2d_batch_destroy(sw2d_batch *b)
{
    if (!b) return SW2D_OK;
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    // members still attached or merely created in the pool go with the batch
    std::vector<sw2d_handle *> ms = b->owned;
    for (sw2d_handle *h : ms) (void)sw2d_destroy(h);
    for (auto &ev : b->events) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    bdrop_graph(b);
    if (b->dtiles) (void)hipFree(b->dtiles);
    if (b->dscen) (void)hipFree(b->dscen);
    if (b->dstatus) (void)hipFree(b->dstatus);
    if (b->hstatus) (void)hipHostFree(b->hstatus);
    if (b->pool_base) (void)hipFree(b->pool_base);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
    return SW2D_OK;
}

const char *sw2d_batch_last_error(const sw2d_batch *b) { return b ? b->err.c_str() : g_create_error.c_str(); }

size_t sw2d_batch_footprint(const sw2d_grid_desc *grid, const sw2d_params *params, int tile_rows,
                            long hydro_cap, int manning_map)
{
    if (!grid || !params || grid->rows < 2 || grid->cols < 2) return 0;
    const size_t R = grid->rows, C = grid->cols;
    const size_t P = (size_t)round_up((int)C, 16);
    const size_t hb = depth_bytes(params->precision), fb = flux_bytes(params->precision);
    const int H = params->tile_rows > 0 ? params->tile_rows : (tile_rows > 0 ? tile_rows : 32);
    const size_t ntiles = ((C + 57) / 58) * ((R + H - 1) / H);
    size_t b = 0;
    b += 2 * pool_bytes_of(R * P * hb) + 4 * pool_bytes_of(R * P * fb);  // h, um, vn ping-pong
    b += pool_bytes_of(R * P * 8);                               // z
    if (params->precision == SW2D_FP32) b += pool_bytes_of(R * P * 8);  // bed steps
    b += pool_bytes_of(sizeof(Scenario));                        // descriptor
    b += pool_bytes_of(3 * sizeof(Clock));
    b += pool_bytes_of(2 * ntiles * sizeof(Partial));
    b += 2 * pool_bytes_of(sizeof(Ledger));                      // ledger, peek
    if (hydro_cap > 0) b += pool_bytes_of(2 * (size_t)hydro_cap * 8);
    if (manning_map) b += pool_bytes_of(R * P);
    b += 8 * kPoolAlign;                                         // forcing tables, sources, outlet
    return b;
}

int sw2d_batch_new(sw2d_batch *b, const sw2d_grid_desc *grid, const double *z, const double *h0,
                   const sw2d_params *params, sw2d_handle **out)
{
    if (!b) return bfail(nullptr, SW2D_E_ARG, "batch is NULL");
    int rc = create_impl(grid, z, h0, params, b, out);
    if (rc) b->err = g_create_error;
    return rc;
}

int sw2d_batch_attach(sw2d_batch *b, sw2d_handle *h, long max_iters, double t_end)
{
    if (!b || !h) return bfail(b, SW2D_E_ARG, "NULL argument");
    if (h->batch != b) return bfail(b, SW2D_E_ARG, "handle does not belong to this batch");
    if (h->attached) return bfail(b, SW2D_E_STATE, "already attached");
    if (b->free_slots.empty()) return bfail(b, SW2D_E_STATE, "batch is full");
    BCHK(b, hipStreamSynchronize(b->stream));
    Clock ck;
    BCHK(b, hipMemcpy(&ck, h->clk + (h->step_index % 3), sizeof(Clock), hipMemcpyDeviceToHost));
    const int slot = b->free_slots.back();
    b->free_slots.pop_back();
    h->member_slot = slot;
    h->phase0 = (int)(((b->phase - h->step_index) % 6 + 6) % 6);
    h->k_end = max_iters < 0 ? 0x7fffffffffffffffLL : h->step_index + max_iters;
    h->t_end = (t_end == t_end) ? t_end : INFINITY;
    Status s;
    memset(&s, 0, sizeof s);
    s.k = h->step_index;
    s.t = ck.t;
    BCHK(b, hipMemcpy(b->dstatus + slot, &s, sizeof s, hipMemcpyHostToDevice));
    h->attached = true;
    b->members.push_back(h);
    b->dirty = true;
    return SW2D_OK;
}

int sw2d_batch_detach(sw2d_batch *b, sw2d_handle *h)
{
    if (!b || !h) return bfail(b, SW2D_E_ARG, "NULL argument");
    if (h->batch != b || !h->attached) return bfail(b, SW2D_E_STATE, "not attached to this batch");
    BCHK(b, hipStreamSynchronize(b->stream));
    b->members.erase(std::find(b->members.begin(), b->members.end(), h));
    b->free_slots.push_back(h->member_slot);
    h->member_slot = -1;
    h->attached = false;
    h->desc_dirty = true;
    b->dirty = true;
    return SW2D_OK;
}

int sw2d_batch_step(sw2d_batch *b, int n_iters, int *running)
{
    if (!b) return SW2D_E_ARG;
    if (n_iters < 0) return bfail(b, SW2D_E_ARG, "negative iteration count");
    int rc = rebuild(b);
    if (rc) return rc;
    if (b->members.empty()) {
        if (running) *running = 0;
        return SW2D_OK;
    }
    long n = n_iters;
    while (n > 0) {
        const bool group = b->graph_on && (b->phase % 6 == 0) && n >= 6;
        if (group && (rc = ensure_bgraph(b))) return rc;
        const long m = group ? 6 * (n / 6) : 1;
        const long reps = group ? m / 6 : 1;
        for (long r = 0; r < reps; r++) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (b->timing) {
                BCHK(b, hipEventCreate(&e0));
                BCHK(b, hipEventCreate(&e1));
                BCHK(b, hipEventRecord(e0, b->stream));
            }
            if (group) {
                BCHK(b, hipGraphLaunch(b->gexec, b->stream));
            } else {
                hipError_t e = blaunch(b, b->dscen, b->dtiles, b->nblocks, (int)(b->phase % 6), b->features);
                if (e != hipSuccess) return bhip(b, e, "batch launch");
            }
            if (b->timing) {
                BCHK(b, hipEventRecord(e1, b->stream));
                b->events.emplace_back(e0, e1);
                b->launches += group ? 6 : 1;
                if (b->events.size() >= 1024 && flush_bevents(b)) return SW2D_E_HIP;
            }
        }
        b->phase += m;
        n -= m;
    }
    return refresh(b, running);
}

int sw2d_batch_status(sw2d_batch *b, const sw2d_handle *h, int *done, long *iters, double *time)
{
    if (!b || !h) return bfail(b, SW2D_E_ARG, "NULL argument");
    if (h->batch != b || !h->attached) return bfail(b, SW2D_E_STATE, "not attached to this batch");
    const Status &s = b->hstatus[h->member_slot];
    if (done) *done = s.done || s.fin;
    if (iters) *iters = (long)s.k;
    if (time) *time = s.t;
    return SW2D_OK;
}

int sw2d_batch_autotune(sw2d_batch *b, int iters, int *tile_rows)
{
    if (!b) return SW2D_E_ARG;
    if (iters < 1) return bfail(b, SW2D_E_ARG, "iters must be positive");
    if (b->members.empty()) return bfail(b, SW2D_E_STATE, "no members to tune on");
    // at any point of the run: the members' progress from the device first
    int rc = rebuild(b);
    if (rc) return rc;
    if ((rc = refresh(b, nullptr))) return rc;
    const int nm = (int)b->members.size();
    std::vector<int> cand = {8, 12, 16, 24, 32, 48, 64, 96, 128};
    // heights whose launch over all members fills whole rounds of wave slots
    int feats_all = 0;
    for (sw2d_handle *h : b->members) feats_all |= features_of(make_common(h));
    const long slots = wave_slots(b->device, b->precision, feats_all);
    for (int H : round_fill_heights(slots, [&](int H) {
             long n = 0;
             for (const sw2d_handle *h : b->members)
                 n += blocks_at(h, h->prm.tile_rows > 0 ? h->prm.tile_rows : H);
             return n;
         }))
        cand.push_back(H);
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    // scratch bookkeeping per member, sized for the smallest candidate
    struct Scratch { Clock *clk; Partial *part; Ledger *led; double *bc; int maxt; };
    std::vector<Scratch> sc(nm);
    Scenario *dscen_s = nullptr;
    TileRef *dtiles_s = nullptr;
    auto cleanup = [&]() {
        for (int i = 0; i < nm; i++) {
            sw2d_handle *h = b->members[i];
            dfree(h, sc[i].clk); dfree(h, sc[i].part); dfree(h, sc[i].led); dfree(h, sc[i].bc);
        }
        if (dscen_s) (void)hipFree(dscen_s);
        if (dtiles_s) (void)hipFree(dtiles_s);
    };
    long maxblocks = 0;
    for (int i = 0; i < nm; i++) {
        sw2d_handle *h = b->members[i];
        sc[i] = Scratch{nullptr, nullptr, nullptr, nullptr, 0};
        // the most tiles a trial gives this member: its own fixed height, or
        // the smallest candidate
        const int hmin = h->prm.tile_rows > 0 ? h->prm.tile_rows : cand[0];
        sc[i].maxt = ((h->C + h->W - 1) / h->W) * ((h->R + hmin - 1) / hmin);
        maxblocks += sc[i].maxt + 1 + (h->nbc + 63) / 64;
    }
    for (int i = 0; i < nm && !rc; i++) {
        sw2d_handle *h = b->members[i];
        if ((rc = dalloc(h, &sc[i].clk, 3)) || (rc = dalloc(h, &sc[i].part, 2 * (size_t)sc[i].maxt)) ||
            (rc = dalloc(h, &sc[i].led, 1)) || (rc = dalloc(h, &sc[i].bc, std::max(h->nbc, 1)))) {
            b->err = h->err;
        }
    }
    if (rc) { cleanup(); return rc; }
    hipError_t e = hipMalloc(&dscen_s, sizeof(Scenario) * nm);
    if (e == hipSuccess) e = hipMalloc(&dtiles_s, sizeof(TileRef) * maxblocks);
    if (e != hipSuccess) { cleanup(); return bhip(b, e, "hipMalloc"); }
    hipEvent_t e0, e1;
    BCHK(b, hipEventCreate(&e0));
    BCHK(b, hipEventCreate(&e1));
    // trial launches run the next iteration of every member still running,
    // at the live launch phase: they read its clock (copied) and live buffer
    // and write only the spare buffer and scratch bookkeeping.  trial(H, n):
    // mean time of n launches after one warm-up (negative when no member runs)
    const int phase = (int)(b->phase % 6);
    auto trial = [&](int H, int n, int order, int cap) -> float {
        std::vector<Scenario> desc(nm);
        std::vector<TileRef> tiles;
        std::vector<int> counts, sids;
        int feats = 0;
        for (int i = 0; i < nm && e == hipSuccess; i++) {
            sw2d_handle *h = b->members[i];
            const Status &st = b->hstatus[h->member_slot];
            if (st.done || st.fin) continue;
            const int H0 = h->H;
            set_tiles(h, h->prm.tile_rows > 0 ? h->prm.tile_rows : H);
            Scenario d = make_scenario(h);
            set_tiles(h, H0);
            d.s.clk = sc[i].clk;
            d.part = sc[i].part;
            d.s.nparts = d.ntiles;
            d.s.ledger = sc[i].led;
            d.s.hydro = nullptr;
            d.s.bc_acc = sc[i].bc;
            d.phase0 = h->phase0;
            feats |= features_of(d.s);
            desc[i] = d;
            counts.push_back(d.nblocks);
            sids.push_back(i);
            e = hipMemcpyAsync(sc[i].clk, h->clk, 3 * sizeof(Clock), hipMemcpyDeviceToDevice, b->stream);
        }
        order_tiles(counts, sids, order, tiles);
        if (e != hipSuccess || tiles.empty()) return -1.0f;
        const size_t lds_live = b->lds_pad;
        b->lds_pad = lds_for_cap(b->device, cap);
        e = hipMemcpy(dscen_s, desc.data(), sizeof(Scenario) * nm, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(dtiles_s, tiles.data(), sizeof(TileRef) * tiles.size(), hipMemcpyHostToDevice);
        if (e == hipSuccess) e = blaunch(b, dscen_s, dtiles_s, (int)tiles.size(), phase, feats);
        if (e == hipSuccess) e = hipEventRecord(e0, b->stream);
        for (int i = 0; i < n && e == hipSuccess; i++) e = blaunch(b, dscen_s, dtiles_s, (int)tiles.size(), phase, feats);
        if (e == hipSuccess) e = hipEventRecord(e1, b->stream);
        if (e == hipSuccess) e = hipEventSynchronize(e1);
        float ms = 0;
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, e0, e1);
        b->lds_pad = lds_live;
        if (e == hipSuccess) b->tune_log.push_back({(double)H, (double)order, (double)cap, ms / n});
        return e == hipSuccess ? ms / n : -1.0f;
    };
    b->tune_log.clear();
    // two stages: a short trial of every candidate, then `iters` launches of
    // the three fastest
    int best = b->H;
    std::vector<std::pair<float, int>> t1;
    for (int H : cand) {
        const float ms = trial(H, cand.size() > 3 ? std::max(1, iters / 4) : iters, b->order,
                               b->waves_cap);
        if (ms < 0) break;
        t1.emplace_back(ms, H);
    }
    std::sort(t1.begin(), t1.end());
    int best_order = b->order, best_cap = b->waves_cap;
    if (e == hipSuccess && !t1.empty()) {
        best = t1.front().second;
        float best_ms = t1.front().first;
        if (t1.size() > 3) {
            best_ms = 1e30f;
            for (size_t i = 0; i < 3 && e == hipSuccess; i++) {
                const float ms = trial(t1[i].second, iters, b->order, b->waves_cap);
                if (ms >= 0 && ms < best_ms) { best_ms = ms; best = t1[i].second; }
            }
        }
        // the launch schedule at that height: block order x a cap on the
        // resident waves per SIMD below what the registers allow; a change is
        // kept only when it is at least 1 % faster
        const int natural = std::max(1, fused_waves_per_cu(b->precision, feats_all) / 4);
        std::vector<std::pair<int, int>> sched = {{0, 0}, {1, 0}};
        for (int cap = natural - 1; cap >= std::max(1, natural - 2); cap--) {
            sched.push_back({0, cap});
            sched.push_back({1, cap});
        }
        const float ref = trial(best, iters, b->order, b->waves_cap);
        float pick_ms = ref;
        for (auto [order, cap] : sched) {
            if (e != hipSuccess) break;
            if (order == b->order && cap == b->waves_cap) continue;
            const float ms = trial(best, iters, order, cap);
            if (ms >= 0 && ms < 0.99f * pick_ms) { pick_ms = ms; best_order = order; best_cap = cap; }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    cleanup();
    if (e != hipSuccess) return bhip(b, e, "batch autotune");
    b->H = best;
    if (best_order != b->order || best_cap != b->waves_cap) {
        b->order = best_order;
        b->waves_cap = best_cap;
        b->lds_pad = lds_for_cap(b->device, best_cap);
        bdrop_graph(b);
    }
    // retiling folds each member's pending partial sums into one record
    for (sw2d_handle *h : b->members) {
        if (h->prm.tile_rows > 0) continue;
        if ((rc = retile(h, best))) { b->err = h->err; return rc; }
    }
    b->dirty = true;
    BCHK(b, hipStreamSynchronize(b->stream));
    if (tile_rows) *tile_rows = best;
    return SW2D_OK;
}

int sw2d_batch_set_schedule(sw2d_batch *b, int order, int waves_per_simd)
{
    if (!b) return SW2D_E_ARG;
    if (order < 0 || order > 1 || waves_per_simd < 0 || waves_per_simd > 8)
        return bfail(b, SW2D_E_ARG, "order must be 0 or 1 and waves_per_simd in [0, 8]");
    BCHK(b, hipStreamSynchronize(b->stream));
    b->order = order;
    b->waves_cap = waves_per_simd;
    b->lds_pad = lds_for_cap(b->device, waves_per_simd);
    bdrop_graph(b);
    b->dirty = true;
    return SW2D_OK;
}

int sw2d_batch_schedule(sw2d_batch *b, int *order, int *waves_per_simd)
{
    if (!b) return SW2D_E_ARG;
    if (order) *order = b->order;
    if (waves_per_simd) *waves_per_simd = b->waves_cap;
    return SW2D_OK;
}

int sw2d_batch_tune_report(sw2d_batch *b, double *rows, int cap, int *n)
{
    if (!b) return SW2D_E_ARG;
    const int m = (int)b->tune_log.size();
    for (int i = 0; i < std::min(m, cap) && rows; i++)
        for (int k = 0; k < 4; k++) rows[4 * i + k] = b->tune_log[i][k];
    if (n) *n = m;
    return SW2D_OK;
}

int sw2d_batch_memory(sw2d_batch *b, size_t *in_use, size_t *peak, size_t *capacity)
{
    if (!b) return SW2D_E_ARG;
    if (in_use) *in_use = b->pool.in_use();
    if (peak) *peak = b->pool.peak();
    if (capacity) *capacity = b->pool.capacity();
    return SW2D_OK;
}

int sw2d_batch_timing(sw2d_batch *b, int enable)
{
    if (!b) return SW2D_E_ARG;
    if (flush_bevents(b)) return SW2D_E_HIP;
    b->kernel_ms = 0.0;
    b->launches = 0;
    b->timing = enable ? 1 : 0;
    return SW2D_OK;
}

int sw2d_batch_kernel_time(sw2d_batch *b, double *ms, long *launches)
{
    if (!b) return SW2D_E_ARG;
    if (flush_bevents(b)) return SW2D_E_HIP;
    if (ms) *ms = b->kernel_ms;
    if (launches) *launches = (long)b->launches;
    return SW2D_OK;
}

}  // extern "C"
