// sw2d_api.hip -- host implementation of the C ABI declared in include/sw2d.h.
#include <math.h>
#include <stdio.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "sw2d_common.hpp"
#include "sw2d_ops.hpp"

using namespace sw2d;

namespace sw2d {

int round_up(int x, int m) { return (x + m - 1) / m * m; }

// the StepCommon of iteration k = h->step_index
StepCommon make_common(sw2d_handle *h)
{
    StepCommon s{};
    const long long k = h->step_index;
    s.R = h->R; s.C = h->C;
    s.dx = h->grid.cellsize; s.dy = h->grid.cellsize;
    s.g = h->prm.g; s.rn = h->prm.manning; s.th = h->prm.th;
    s.P_lake = 1.00 - h->prm.LWL;
    s.P_land = 1.00 - (h->prm.INT + h->prm.INF);
    s.EV_WL_min = h->prm.EV_WL_min;
    s.fixed_dt = !h->prm.adaptive;
    s.dt_fixed = h->prm.dt;
    s.dt2_fixed = 2.00 * h->prm.dt;
    s.courant = h->prm.courant;
    s.dt2_max = 2.00 * h->prm.dt;
    s.clk = h->clk;
    s.slot_a = (int)(k % 3); s.slot_b = (int)((k + 1) % 3); s.slot_c = (int)((k + 2) % 3);
    s.step_index = k;
    s.rain = h->rain; s.nrain = h->nrain; s.dtrain = h->dtrain;
    s.evapo = h->evapo; s.nevapo = h->nevapo; s.evap_on = h->evap_on;
    s.Q = h->Q; s.nq = h->nq; s.nst = h->nst; s.dtoq = h->dtoq;
    s.src_r = h->src_r; s.src_c = h->src_c;
    s.storm_on = h->storm_on;
    s.st_peak = h->storm.peak; s.st_x0 = h->storm.x0; s.st_y0 = h->storm.y0;
    s.st_vx = h->storm.vx; s.st_vy = h->storm.vy; s.st_sigma = h->storm.sigma;
    s.st_t0 = h->storm.t0; s.st_t1 = h->storm.t1;
    s.ncls_on = h->ncls_on;
    for (int i = 0; i < kMaxClasses; i++) s.ntab[i] = h->ntab[i];
    s.out_r = h->out_r; s.out_c = h->out_c; s.out_edge = h->out_edge;
    s.part_out = h->part + (size_t)(k % 2) * h->nparts;
    s.part_in = h->part + (size_t)((k + 1) % 2) * h->nparts;
    s.nparts = h->nparts;
    s.ledger = h->ledger;
    s.hydro = h->hydro; s.hydro_cap = h->hydro_cap;
    s.nbc = h->nbc; s.bc_r = h->bc_r; s.bc_c = h->bc_c; s.bc_kind = h->bc_kind; s.bc_acc = h->bc_acc;
    return s;
}

bool fused(const sw2d_handle *h) { return h->prm.kernel == SW2D_KERNEL_FUSED; }

int fused_blocks(const sw2d_handle *h) { return h->ntiles + 1 + (h->nbc + 63) / 64; }

// the fused kernel's descriptor of a stand-alone handle (phase 0 = iteration 0)
Scenario make_scenario(sw2d_handle *h)
{
    Scenario d;
    memset(&d, 0, sizeof d);
    d.s = make_common(h);
    h->features = features_of(d.s);
    for (int b = 0; b < 2; b++) {
        d.st[b][0] = h->h[b];
        d.st[b][1] = h->um[b];
        d.st[b][2] = h->vn[b];
    }
    d.z = (const double *)h->z;
    d.dz = (const double *)h->dz;
    d.cls = h->cls;
    d.P = h->P; d.W = h->W; d.H = h->H;
    d.nstrips = h->nstrips; d.ntiles = h->ntiles;
    d.nblocks = fused_blocks(h);
    d.part = h->part;
    d.s.nparts = h->nparts;
    d.phase0 = 0;
    d.k_end = 0x7fffffffffffffffLL;
    d.t_end = INFINITY;
    d.status = nullptr;
    return d;
}

int upload_desc(sw2d_handle *h)
{
    if (!h->desc_dirty) return SW2D_OK;
    const Scenario d = make_scenario(h);
    HIPCHK(h, hipMemcpyAsync(h->dscen, &d, sizeof d, hipMemcpyHostToDevice, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->desc_dirty = false;
    return SW2D_OK;
}

hipError_t launch_one(sw2d_handle *h, const Scenario *scen, int phase)
{
    return launch_fused(h->prm.precision, h->stream, scen, nullptr, fused_blocks(h), phase,
                        h->features);
}

// the graph of iterations k .. k+5 (k % 6 == 0): six launches, phases 0..5
int ensure_graph(sw2d_handle *h)
{
    if (h->gexec) return SW2D_OK;
    hipGraph_t g = nullptr;
    HIPCHK(h, hipStreamBeginCapture(h->stream, hipStreamCaptureModeThreadLocal));
    hipError_t e = hipSuccess;
    for (int ph = 0; ph < 6 && e == hipSuccess; ph++) e = launch_one(h, h->dscen, ph);
    hipError_t e2 = hipStreamEndCapture(h->stream, &g);
    if (e != hipSuccess) { if (g) (void)hipGraphDestroy(g); return hip_fail(h, e, "graph capture launch"); }
    if (e2 != hipSuccess) return hip_fail(h, e2, "hipStreamEndCapture");
    e = hipGraphInstantiate(&h->gexec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) { h->gexec = nullptr; return hip_fail(h, e, "hipGraphInstantiate"); }
    return SW2D_OK;
}

// host mirror after m iterations
void advance(sw2d_handle *h, long m)
{
    for (long i = 0; i < m; i++) {
        if (!h->prm.adaptive) {
            h->host_time = h->host_time + h->prm.dt;
            h->host_time = h->host_time + h->prm.dt;
        }
    }
    h->step_index += m;
    h->host_mstep += 2 * m;
    h->cur = (int)(h->step_index & 1);
}

// Switch a fused handle to H rows per tile.  The partial sums the last
// iteration wrote are still to be reduced by the next one; they are folded
// into one record of the new buffer (sized for the new tiling exactly), so no
// record of the old tiling is ever reduced twice.
int retile(sw2d_handle *h, int H)
{
    const int old_n = h->nparts;
    const int H0 = h->H;
    const int new_n = ((h->C + h->W - 1) / h->W) * ((h->R + H - 1) / H);
    if (new_n == old_n && H == H0) return SW2D_OK;
    // the new partial buffer is filled first; the handle changes only once
    // every step has succeeded (a failure leaves tiling and buffer as they were)
    HIPCHK(h, hipStreamSynchronize(h->stream));
    Partial *np = nullptr;
    int rc = dalloc(h, &np, 2 * (size_t)new_n);
    if (rc) return rc;
    if (h->step_index > 0) {
        const size_t pend = (size_t)((h->step_index + 1) % 2);
        std::vector<Partial> old((size_t)old_n);
        hipError_t e = hipMemcpy(old.data(), h->part + pend * old_n, sizeof(Partial) * old_n,
                                 hipMemcpyDeviceToHost);
        Partial sum{0.0, 0.0, 0.0, 0.0};
        for (const Partial &p : old) {
            sum.vol += p.vol; sum.src += p.src; sum.clip += p.clip; sum.qout += p.qout;
        }
        if (e == hipSuccess) e = hipMemcpy(np + pend * new_n, &sum, sizeof sum, hipMemcpyHostToDevice);
        if (e != hipSuccess) {
            dfree(h, np);
            return hip_fail(h, e, "retile");
        }
    }
    set_tiles(h, H);
    dfree(h, h->part);
    h->part = np;
    h->nparts = new_n;
    changed(h);
    return SW2D_OK;
}

int flush_events(sw2d_handle *h)
{
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (auto &ev : h->events) {
        float ms = 0;
        HIPCHK(h, hipEventElapsedTime(&ms, ev.first, ev.second));
        h->kernel_ms += ms;
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    h->events.clear();
    return SW2D_OK;
}

// Resident one-wave blocks of the fused kernel over the device (waves per CU
// at its register use x CUs): one "round" of a launch.
long wave_slots(int device, int prec, int features)
{
    const int per_cu = fused_waves_per_cu(prec, features);
    hipDeviceProp_t prop;
    int cus = 0;
    if (hipGetDeviceProperties(&prop, device) == hipSuccess) cus = prop.multiProcessorCount;
    return (long)std::max(per_cu, 1) * std::max(cus, 1);
}

// blocks of one launch of a stand-alone grid at tile height H
int blocks_at(const sw2d_handle *h, int H)
{
    const int nstrips = (h->C + h->W - 1) / h->W;
    return nstrips * ((h->R + H - 1) / H) + 1 + (h->nbc + 63) / 64;
}

// Tile heights in [2, 128] that fill whole rounds of wave slots: for k = 1..6
// the smallest H whose launch fits in k rounds.  A launch slightly over a
// whole number of rounds leaves a nearly empty last round (on a 2048^2 grid
// in FP64, 2016 tiles at H = 37 run 12% faster than 2 rounds at H = 20 and
// 17% faster than 2.25 rounds at H = 16), and within a round a larger H
// re-reads fewer halo rows per output row.  On grids too small to fill one
// round the launch time is the tile's row chain (H + 6 rows in sequence), so
// the smallest height that still fits one round is the fastest (256^2: 16.7
// us per iteration at H = 2 against 28.2 at H = 8).
std::vector<int> round_fill_heights(long slots, const std::function<long(int)> &blocks)
{
    std::vector<int> out;
    int H = 2;
    for (int k = 6; k >= 1; k--) {   // the smallest height grows as k falls
        while (H <= 128 && blocks(H) > k * slots) H++;
        if (H > 128) break;
        if (out.empty() || out.back() != H) out.push_back(H);
    }
    return out;   // ascending
}

// Default tile height of the fused kernel (every tile is one wave): the
// smallest height that fills whole rounds of wave slots, taking the fewest
// rounds whose height is at most 48 (16 when none is).  sw2d_autotune replaces
// it by a measured choice.
int choose_tile_rows(const sw2d_handle *h)
{
    if (h->prm.tile_rows > 0) return std::min(h->prm.tile_rows, 512);
    const long slots = wave_slots(h->device, h->prm.precision,
                                  features_of(make_common(const_cast<sw2d_handle *>(h))));
    const std::vector<int> fit = round_fill_heights(slots, [&](int H) { return (long)blocks_at(h, H); });
    for (auto it = fit.rbegin(); it != fit.rend(); ++it)
        if (*it <= 48) return *it;
    return 16;
}

void set_tiles(sw2d_handle *h, int H)
{
    if (H != h->H) changed(h);
    h->H = H;
    h->nstrips = (h->C + h->W - 1) / h->W;
    h->nchunks = (h->R + H - 1) / H;
    h->ntiles = h->nstrips * h->nchunks;
}

int upload_source_cells(sw2d_handle *h, const int *rr, const int *cc, int n)
{
    dfree(h, h->src_r); dfree(h, h->src_c);
    h->src_r = h->src_c = nullptr;
    int rc;
    if ((rc = dalloc(h, &h->src_r, n)) || (rc = dalloc(h, &h->src_c, n))) return rc;
    if (n > 0) {
        HIPCHK(h, hipMemcpy(h->src_r, rr, sizeof(int) * n, hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->src_c, cc, sizeof(int) * n, hipMemcpyHostToDevice));
    }
    return SW2D_OK;
}

int check_idle(sw2d_handle *h)
{
    if (!h) return SW2D_E_ARG;
    return SW2D_OK;
}

}  // namespace

extern "C" {

void sw2d_default_params(sw2d_params *p)
{
    if (!p) return;
    memset(p, 0, sizeof(*p));
    p->g = 9.80;
    p->manning = 0.03;
    p->th = 1.0e-4;
    p->dt = 0.5;
    p->precision = SW2D_FP64;
    p->kernel = SW2D_KERNEL_FUSED;
    p->courant = 0.5;
}

int sw2d_version(void) { return SW2D_ABI_VERSION; }

const char *sw2d_last_error(const sw2d_handle *h)
{
    return h ? h->err.c_str() : g_create_error.c_str();
}

}  // extern "C"

namespace sw2d {

void prepare_images(const sw2d_grid_desc *grid, const double *z, const double *h0, int precision,
                    HostImages &img)
{
    const int R = grid->rows, C = grid->cols;
    const size_t RC = (size_t)R * C;
    const double nodata = grid->nodata;
    double v0 = 0.0;
    for (size_t i = 0; i < RC; i++)
        if (z[i] != nodata) v0 += h0[i];
    img.volume0 = v0 * grid->cellsize * grid->cellsize;
    img.P = round_up(C, 16);
    const size_t n = (size_t)R * img.P;
    img.z.assign(n, 0.0);
    for (int r = 0; r < R; r++)
        for (int j = 0; j < C; j++) {
            const double zv = z[(size_t)r * C + j];
            img.z[(size_t)r * img.P + j] = (zv != nodata) ? zv : NAN;
        }
    img.h64.clear();
    img.h32.clear();
    if (depth_bytes(precision) == 8) {
        img.h64.assign(n, 0.0);
        for (int r = 0; r < R; r++)
            for (int j = 0; j < C; j++) img.h64[(size_t)r * img.P + j] = h0[(size_t)r * C + j];
    } else {
        img.h32.assign(n, 0.0f);
        for (int r = 0; r < R; r++)
            for (int j = 0; j < C; j++) img.h32[(size_t)r * img.P + j] = (float)h0[(size_t)r * C + j];
    }
}

int create_impl(const sw2d_grid_desc *grid, const double *z, const double *h0,
                const sw2d_params *params, sw2d_batch *batch, sw2d_handle **out, const HostImages *img)
{
    if (!out) return fail(nullptr, SW2D_E_ARG, "out is NULL");
    *out = nullptr;
    if (!grid || (!img && (!z || !h0)) || !params) return fail(nullptr, SW2D_E_ARG, "NULL argument");
    if (grid->rows < 2 || grid->cols < 2) return fail(nullptr, SW2D_E_ARG, "grid must be at least 2x2");
    if (!(grid->cellsize > 0)) return fail(nullptr, SW2D_E_ARG, "cellsize must be positive");
    if (!known_precision(params->precision)) return fail(nullptr, SW2D_E_ARG, "unknown precision");
    if (params->kernel != SW2D_KERNEL_FUSED && params->kernel != SW2D_KERNEL_PLAIN)
        return fail(nullptr, SW2D_E_ARG, "unknown kernel");
    if (params->kernel == SW2D_KERNEL_PLAIN && params->precision != SW2D_FP64)
        return fail(nullptr, SW2D_E_UNSUPPORTED, "the plain kernel sequence is FP64 only");
    if (!(params->dt > 0)) return fail(nullptr, SW2D_E_ARG, "dt must be positive");
    if (params->adaptive && !(params->courant > 0)) return fail(nullptr, SW2D_E_ARG, "courant must be positive");
    if (params->tile_rows < 0 || params->tile_rows > 512)
        return fail(nullptr, SW2D_E_ARG, "tile_rows must be in [0, 512]");
    {
        // offsets in the kernels are 32-bit: byte offsets of doubles in the
        // padded grid (pitch a multiple of 16; the fused kernel addresses its
        // arrays through buffer resources) and element offsets of the plain
        // layout (R*C + C/2 plus 2C + 4 of tail) must fit, so both stay below
        // 2^29 cells
        const long long R = grid->rows, C = grid->cols;
        const long long padded = R * ((C + 15) / 16 * 16), flat = R * C + C / 2 + 2 * C + 4;
        if (padded >= (1LL << 29) || flat >= (1LL << 29))
            return fail(nullptr, SW2D_E_ARG, "grid too large: rows * cols must stay below 2^29 cells");
    }
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0)
        return fail(nullptr, SW2D_E_NODEV, "no HIP device");
    if (params->device < 0 || params->device >= ndev) return fail(nullptr, SW2D_E_NODEV, "bad device ordinal");

    if (batch) {
        if (params->kernel != SW2D_KERNEL_FUSED)
            return fail(nullptr, SW2D_E_UNSUPPORTED, "batch members use the fused kernel");
        if (params->precision != batch->precision)
            return fail(nullptr, SW2D_E_ARG, "member precision differs from the batch's");
    }
    if (img && (params->kernel != SW2D_KERNEL_FUSED || img->P != round_up(grid->cols, 16) ||
                img->z.size() != (size_t)grid->rows * img->P ||
                (depth_bytes(params->precision) == 8 ? img->h64.size() : img->h32.size()) != img->z.size()))
        return fail(nullptr, SW2D_E_ARG, "prepared images do not match the grid and precision");

    sw2d_handle *h = new sw2d_handle();
    h->graph_on = sw2d::graphs_enabled();
    h->batch = batch;
    h->grid = *grid;
    h->prm = *params;
    h->R = grid->rows; h->C = grid->cols;
    h->device = batch ? batch->device : params->device;
    h->prm.device = h->device;
    int rc = SW2D_OK;
    auto bail = [&](int code) {
        g_create_error = h->err;
        sw2d_destroy(h);
        return code;
    };
    hipError_t e = hipSetDevice(h->device);
    if (e != hipSuccess) { hip_fail(h, e, "hipSetDevice"); return bail(SW2D_E_HIP); }
    if (batch) {
        h->stream = batch->stream;
    } else {
        e = hipStreamCreate(&h->stream);  // blocking stream: ordered with hipMemcpy on the null stream
        if (e != hipSuccess) { hip_fail(h, e, "hipStreamCreate"); return bail(SW2D_E_HIP); }
    }

    const int R = h->R, C = h->C;
    const size_t RC = (size_t)R * C;

    if (fused(h)) {
        HostImages own;
        if (!img) {
            try {
                prepare_images(grid, z, h0, params->precision, own);
            } catch (const std::bad_alloc &) {
                fail(h, SW2D_E_NOMEM, "host memory for the input images");
                return bail(SW2D_E_NOMEM);
            }
            img = &own;
        }
        h->volume0 = img->volume0;
        h->P = round_up(C, 16);
        const size_t n = (size_t)R * h->P;
        h->W = 58;
        set_tiles(h, (batch && batch->H > 0 && params->tile_rows == 0) ? batch->H : choose_tile_rows(h));
        h->nparts = h->ntiles;
        // depths in the form's depth type, fluxes in its flux type
        const int hb = depth_bytes(params->precision), fb = flux_bytes(params->precision);
        void **bufs[] = {&h->h[0], &h->h[1], &h->um[0], &h->um[1], &h->vn[0], &h->vn[1]};
        for (int i = 0; i < 6; i++) {
            void **b = bufs[i];
            const int eb = i < 2 ? hb : fb;
            if (eb == 8) { if ((rc = dalloc(h, (double **)b, n))) return bail(rc); }
            else if ((rc = dalloc(h, (float **)b, n))) return bail(rc);
        }
        if ((rc = dalloc(h, (double **)&h->z, n))) return bail(rc);
        // the FP32 form's kernel reads its bed steps (8 bytes a cell, as z)
        if (params->precision == SW2D_FP32 && (rc = dalloc(h, (double **)&h->dz, n))) return bail(rc);
        if ((rc = dalloc(h, &h->dscen, 1))) return bail(rc);
        // pitched images: inactive cells carry NaN in z, which stays FP64
        e = hipMemcpy(h->z, img->z.data(), n * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess)
            e = hb == 8 ? hipMemcpy(h->h[0], img->h64.data(), n * 8, hipMemcpyHostToDevice)
                        : hipMemcpy(h->h[0], img->h32.data(), n * 4, hipMemcpyHostToDevice);
        if (e == hipSuccess && h->dz) e = launch_bed_steps(h);
        if (e != hipSuccess) { hip_fail(h, e, "upload"); return bail(SW2D_E_HIP); }
    } else {
        const double nodata = grid->nodata;
        double v0 = 0.0;  // initial volume
        for (size_t i = 0; i < RC; i++)
            if (z[i] != nodata) v0 += h0[i];
        h->volume0 = v0 * grid->cellsize * grid->cellsize;
        h->N = R * C + C / 2;
        h->NA = h->N + 2 * C + 4;
        double **arrs[] = {&h->pz, &h->ph, &h->pho, &h->phm, &h->phn, &h->pum, &h->pumo, &h->puu1,
                           &h->pvv1, &h->puua, &h->pvva, &h->pvn, &h->pvno, &h->puu, &h->pvv};
        for (double **a : arrs)
            if ((rc = dalloc(h, a, h->NA))) return bail(rc);
        if ((rc = dalloc(h, &h->pinf, h->NA))) return bail(rc);
        std::vector<int> inf;
        try {
            inf.assign(h->NA, 0);
        } catch (const std::bad_alloc &) {
            fail(h, SW2D_E_NOMEM, "host memory for the activity mask");
            return bail(SW2D_E_NOMEM);
        }
        for (size_t i = 0; i < RC; i++) inf[i] = (z[i] != nodata) ? 1 : 0;
        e = hipMemcpy(h->pz, z, RC * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h->ph, h0, RC * 8, hipMemcpyHostToDevice);
        if (e == hipSuccess) e = hipMemcpy(h->pinf, inf.data(), h->NA * sizeof(int), hipMemcpyHostToDevice);
        if (e != hipSuccess) { hip_fail(h, e, "upload"); return bail(SW2D_E_HIP); }
        int blocks = (h->N + 255) / 256;
        h->plain_blocks = std::min(blocks, 2048);
        h->nparts = h->plain_blocks;
    }

    if ((rc = dalloc(h, &h->clk, 3)) || (rc = dalloc(h, &h->part, 2 * (size_t)h->nparts)) ||
        (rc = dalloc(h, &h->ledger, 1)) || (rc = dalloc(h, &h->peek, 1)))
        return bail(rc);
    Clock ck[3];
    memset(ck, 0, sizeof(ck));
    for (auto &c : ck) c.cfl_bits = kInfBits;
    ck[0].t = params->time0;
    e = hipMemcpy(h->clk, ck, sizeof(ck), hipMemcpyHostToDevice);
    if (e != hipSuccess) { hip_fail(h, e, "clock upload"); return bail(SW2D_E_HIP); }
    h->host_time = params->time0;
    h->host_mstep = 0;
    h->step_index = 0;

    e = !fused(h) ? launch_plain_init(h) : launch_init(h);
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e != hipSuccess) { hip_fail(h, e, "init kernel"); return bail(SW2D_E_HIP); }
    if (batch) batch->owned.push_back(h);
    *out = h;
    return SW2D_OK;
}

}  // namespace sw2d

extern "C" {

int sw2d_create(const sw2d_grid_desc *grid, const double *z, const double *h0,
                const sw2d_params *params, sw2d_handle **out)
{
    return create_impl(grid, z, h0, params, nullptr, out);
}

int sw2d_destroy(sw2d_handle *h)
{
    if (!h) return SW2D_OK;
    if (h->attached) (void)sw2d_batch_detach(h->batch, h);
    if (h->batch) {
        auto &own = h->batch->owned;
        auto it = std::find(own.begin(), own.end(), h);
        if (it != own.end()) own.erase(it);
    }
    if (h->stream) (void)hipStreamSynchronize(h->stream);
    snapshot_free(h);
    for (auto &ev : h->events) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    drop_graph(h);
    for (void *p : h->allocs) {
        if (h->batch) (void)h->batch->pool.release(p);
        else (void)hipFree(p);
    }
    h->allocs.clear();
    if (h->stream && !h->batch) (void)hipStreamDestroy(h->stream);
    delete h;
    return SW2D_OK;
}

int sw2d_set_rain_series(sw2d_handle *h, const double *rain_mm, int n, double dtrain)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (n < 0 || (n > 0 && !rain_mm)) return fail(h, SW2D_E_ARG, "bad rain series");
    if (!(dtrain > 0)) return fail(h, SW2D_E_ARG, "dtrain must be positive");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    dfree(h, h->rain);
    h->rain = nullptr;
    int rc = dalloc(h, &h->rain, n);
    if (rc) return rc;
    if (n > 0) HIPCHK(h, hipMemcpy(h->rain, rain_mm, sizeof(double) * n, hipMemcpyHostToDevice));
    h->nrain = n;
    h->dtrain = dtrain;
    changed(h);
    return SW2D_OK;
}

int sw2d_set_sources(sw2d_handle *h, const double *x, const double *y, int nst, const double *Q,
                     int nq, double dtoq)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (nst < 0 || nq < 0 || (nst > 0 && (!x || !y || (nq > 0 && !Q))))
        return fail(h, SW2D_E_ARG, "bad source arguments");
    if (nst > 0 && !(dtoq > 0)) return fail(h, SW2D_E_ARG, "dtoq must be positive");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    // cell index of stream_flow (SW2D-EFDC.cu:541-544)
    const double res = h->grid.cellsize;
    const double xcoor = h->grid.xllcorner / res, ycoor = h->grid.yllcorner / res;
    std::vector<int> rr(nst), cc(nst);
    for (int i = 0; i < nst; i++) {
        const double brx = x[i] / res, bry = y[i] / res;
        const int qinx = (int)round(fabs(xcoor - brx));
        const int qiny = h->R - (int)round(fabs(ycoor - bry));
        const long idx = (long)(qiny + 1) * h->C - (h->C - (qinx + 1));
        if (idx < 0 || idx >= (long)h->R * h->C) {
            char m[160];
            snprintf(m, sizeof m, "source %d at (%g, %g) maps to cell %ld outside the grid", i, x[i], y[i], idx);
            return fail(h, SW2D_E_ARG, m);
        }
        rr[i] = (int)(idx / h->C);
        cc[i] = (int)(idx % h->C);
    }
    int rc = upload_source_cells(h, rr.data(), cc.data(), nst);
    if (rc) return rc;
    dfree(h, h->Q);
    h->Q = nullptr;
    if ((rc = dalloc(h, &h->Q, (size_t)nq * nst))) return rc;
    if ((size_t)nq * nst > 0)
        HIPCHK(h, hipMemcpy(h->Q, Q, sizeof(double) * nq * nst, hipMemcpyHostToDevice));
    h->nst = nst; h->nq = nq; h->dtoq = dtoq;
    changed(h);
    return SW2D_OK;
}

int sw2d_set_evaporation(sw2d_handle *h, const double *evapo_mm, int n)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (n < 0 || (n > 0 && !evapo_mm)) return fail(h, SW2D_E_ARG, "bad evaporation series");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    dfree(h, h->evapo);
    h->evapo = nullptr;
    int rc = dalloc(h, &h->evapo, n);
    if (rc) return rc;
    if (n > 0) HIPCHK(h, hipMemcpy(h->evapo, evapo_mm, sizeof(double) * n, hipMemcpyHostToDevice));
    h->nevapo = n;
    h->evap_on = 1;
    changed(h);
    return SW2D_OK;
}

int sw2d_set_manning_map(sw2d_handle *h, const unsigned char *cls, const double *table, int nclass)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (!cls || !table || nclass < 1 || nclass > kMaxClasses)
        return fail(h, SW2D_E_ARG, "bad Manning map");
    const size_t RC = (size_t)h->R * h->C;
    for (size_t i = 0; i < RC; i++)
        if (cls[i] >= nclass) return fail(h, SW2D_E_ARG, "Manning class out of range");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (int i = 0; i < kMaxClasses; i++) h->ntab[i] = i < nclass ? table[i] : 0.0;
    int rc;
    if (fused(h)) {
        if (!h->cls && (rc = dalloc(h, &h->cls, (size_t)h->R * h->P))) return rc;
        std::vector<uint8_t> img((size_t)h->R * h->P, 0);
        for (int r = 0; r < h->R; r++)
            for (int j = 0; j < h->C; j++) img[(size_t)r * h->P + j] = cls[(size_t)r * h->C + j];
        HIPCHK(h, hipMemcpy(h->cls, img.data(), img.size(), hipMemcpyHostToDevice));
    } else {
        if (!h->pncell && (rc = dalloc(h, &h->pncell, h->NA))) return rc;
        std::vector<double> nc(h->NA, 0.0);
        for (size_t i = 0; i < RC; i++) nc[i] = table[cls[i]];
        HIPCHK(h, hipMemcpy(h->pncell, nc.data(), nc.size() * 8, hipMemcpyHostToDevice));
    }
    h->ncls_on = 1;
    changed(h);
    return SW2D_OK;
}

int sw2d_set_storm(sw2d_handle *h, const sw2d_storm *storm)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (!storm || !(storm->sigma > 0)) return fail(h, SW2D_E_ARG, "bad storm");
    if (fused(h) && h->H > 512) return fail(h, SW2D_E_ARG, "tile too tall for the storm table");
    h->storm = *storm;
    h->storm_on = 1;
    changed(h);
    return SW2D_OK;
}

int sw2d_set_outlet(sw2d_handle *h, int row, int col)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (row < 0 || row >= h->R || col < 0 || col >= h->C) return fail(h, SW2D_E_ARG, "outlet outside the grid");
    int edge = 0;
    if (row == h->R - 1) edge = 1;
    else if (row == 0) edge = 2;
    else if (col == 0) edge = 3;
    else if (col == h->C - 1) edge = 4;
    if (!edge) return fail(h, SW2D_E_ARG, "outlet must be on a domain edge");
    h->out_r = row; h->out_c = col; h->out_edge = edge;
    changed(h);
    return SW2D_OK;
}

int sw2d_set_hydrograph(sw2d_handle *h, long cap)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (cap < 0) return fail(h, SW2D_E_ARG, "negative capacity");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    dfree(h, h->hydro);
    h->hydro = nullptr;
    int rc = dalloc(h, &h->hydro, 2 * (size_t)cap);
    if (rc) return rc;
    h->hydro_cap = cap;
    changed(h);
    return SW2D_OK;
}

int sw2d_set_lake_bc(sw2d_handle *h, const int *mask)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (!mask) return fail(h, SW2D_E_ARG, "mask is NULL");
    const int R = h->R, C = h->C;
    const int cap = R * C;
    std::vector<int> cells(4 * (size_t)cap), I(4 * (size_t)cap), J(4 * (size_t)cap), counts(4);
    std::vector<double> q(4 * (size_t)cap);
    std::vector<double> zeros((size_t)R * C, 0.0);
    int rc = sw2d_couple_bc(R, C, h->grid.cellsize, h->grid.cellsize, mask, zeros.data(),
                            zeros.data(), zeros.data(), cap, cells.data(), I.data(), J.data(),
                            q.data(), counts.data());
    if (rc) return fail(h, rc, "coupling scan failed");
    std::vector<int> br, bc, bk;
    for (int k = 0; k < 4; k++)
        for (int i = 0; i < counts[k]; i++) {
            const int p = cells[(size_t)k * cap + i];
            br.push_back(p / C); bc.push_back(p % C); bk.push_back(k);
        }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    dfree(h, h->bc_r); dfree(h, h->bc_c); dfree(h, h->bc_kind); dfree(h, h->bc_acc);
    h->bc_r = h->bc_c = h->bc_kind = nullptr;
    h->bc_acc = nullptr;
    const int n = (int)br.size();
    if ((rc = dalloc(h, &h->bc_r, n)) || (rc = dalloc(h, &h->bc_c, n)) ||
        (rc = dalloc(h, &h->bc_kind, n)) || (rc = dalloc(h, &h->bc_acc, n)))
        return rc;
    if (n > 0) {
        HIPCHK(h, hipMemcpy(h->bc_r, br.data(), n * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->bc_c, bc.data(), n * sizeof(int), hipMemcpyHostToDevice));
        HIPCHK(h, hipMemcpy(h->bc_kind, bk.data(), n * sizeof(int), hipMemcpyHostToDevice));
    }
    HIPCHK(h, hipStreamSynchronize(h->stream));
    h->nbc = n;
    changed(h);
    return SW2D_OK;
}

}  // extern "C"

namespace sw2d {

// the heights that fill whole rounds of wave slots on this grid
static std::vector<int> fill_heights(const sw2d_handle *h)
{
    return round_fill_heights(wave_slots(h->device, h->prm.precision, h->features),
                              [&](int H) { return (long)blocks_at(h, H); });
}

// Times `iters` trial launches of the next iteration at each candidate
// height and retiles to the fastest (see sw2d_autotune)
static int autotune_over(sw2d_handle *h, int iters, std::vector<int> cand, int *tile_rows)
{
    HIPCHK(h, hipStreamSynchronize(h->stream));
    std::sort(cand.begin(), cand.end());
    cand.erase(std::unique(cand.begin(), cand.end()), cand.end());
    // scratch bookkeeping: trials read the real state and write only the spare
    // buffer, a copy of the clock ring, scratch partials/ledger/inflow sums
    // (sized for the smallest height tried: the most tiles)
    const int hmin = *std::min_element(cand.begin(), cand.end());
    const int max_tiles = ((h->C + h->W - 1) / h->W) * ((h->R + hmin - 1) / hmin);
    Clock *clk_s = nullptr;
    Partial *part_s = nullptr;
    Ledger *led_s = nullptr;
    double *bc_s = nullptr;
    Scenario *desc_s = nullptr;
    int rc;
    if ((rc = dalloc(h, &clk_s, 3)) || (rc = dalloc(h, &part_s, 2 * (size_t)max_tiles)) ||
        (rc = dalloc(h, &led_s, 1)) || (rc = dalloc(h, &bc_s, std::max(h->nbc, 1))) ||
        (rc = dalloc(h, &desc_s, 1))) {
        dfree(h, clk_s); dfree(h, part_s); dfree(h, led_s); dfree(h, bc_s); dfree(h, desc_s);
        return rc;
    }
    hipEvent_t e0 = nullptr, e1 = nullptr;
    hipError_t e = hipEventCreate(&e0);
    if (e == hipSuccess) e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        if (e0) (void)hipEventDestroy(e0);
        dfree(h, clk_s); dfree(h, part_s); dfree(h, led_s); dfree(h, bc_s); dfree(h, desc_s);
        return hip_fail(h, e, "hipEventCreate");
    }
    const int H0 = h->H;
    // mean time of n trial launches at height H (after one warm-up launch)
    auto trial = [&](int H, int n) -> float {
        set_tiles(h, H);
        Scenario d = make_scenario(h);
        d.s.clk = clk_s;
        d.part = part_s;
        d.s.nparts = h->ntiles;
        d.s.ledger = led_s;
        d.s.hydro = nullptr;
        d.s.bc_acc = bc_s;
        if ((e = hipMemcpyAsync(clk_s, h->clk, 3 * sizeof(Clock), hipMemcpyDeviceToDevice, h->stream)) ||
            (e = hipMemcpyAsync(desc_s, &d, sizeof d, hipMemcpyHostToDevice, h->stream)) ||
            (e = hipStreamSynchronize(h->stream)))
            return 0.0f;
        // every trial launch is the next iteration k again: it reads clock
        // slot k%3 of the scratch copy and the live buffer k%2, and writes
        // only scratch bookkeeping and the spare buffer (k+1)%2
        const int phase = (int)(h->step_index % 6);
        if ((e = launch_one(h, desc_s, phase)) || (e = hipEventRecord(e0, h->stream))) return 0.0f;
        for (int i = 0; i < n; i++)
            if ((e = launch_one(h, desc_s, phase))) return 0.0f;
        float ms = 0;
        if ((e = hipEventRecord(e1, h->stream)) || (e = hipEventSynchronize(e1)) ||
            (e = hipEventElapsedTime(&ms, e0, e1)))
            return 0.0f;
        return ms / n;
    };
    // two stages: a short trial of every candidate, then `iters` launches of
    // the three fastest (near-equal heights are separated by the longer trial)
    std::vector<std::pair<float, int>> t1;
    for (int H : cand) {
        if (H > h->R && H != cand[0]) break;
        const float ms = trial(H, cand.size() > 3 ? std::max(1, iters / 4) : iters);
        if (e != hipSuccess) break;
        t1.emplace_back(ms, H);
    }
    std::sort(t1.begin(), t1.end());
    int best = H0;
    if (e == hipSuccess && !t1.empty()) {
        best = t1.front().second;
        if (t1.size() > 3) {
            float best_ms = 1e30f;
            for (size_t i = 0; i < 3 && e == hipSuccess; i++) {
                const float ms = trial(t1[i].second, iters);
                if (e == hipSuccess && ms < best_ms) { best_ms = ms; best = t1[i].second; }
            }
        }
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    // trials change the tiling of the host mirror: back to the live one
    // before anything else, also on failure
    set_tiles(h, H0);
    if (e != hipSuccess) (void)hipStreamSynchronize(h->stream);
    dfree(h, clk_s); dfree(h, part_s); dfree(h, led_s); dfree(h, bc_s); dfree(h, desc_s);
    if (e != hipSuccess) return hip_fail(h, e, "autotune launch");
    if ((rc = retile(h, best))) return rc;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (tile_rows) *tile_rows = best;
    return SW2D_OK;
}

}  // namespace sw2d

extern "C" {

int sw2d_autotune(sw2d_handle *h, int iters, int *tile_rows)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (iters < 1) return fail(h, SW2D_E_ARG, "iters must be positive");
    if (h->attached) return fail(h, SW2D_E_STATE, "attached batch members are tuned by sw2d_batch_autotune");
    if (!fused(h)) {
        if (tile_rows) *tile_rows = 0;
        return SW2D_OK;
    }
    std::vector<int> cand = {8, 10, 12, 14, 16, 18, 20, 24, 28, 32, 40, 48, 64, 96, 128};
    for (int H : fill_heights(h)) cand.push_back(H);
    return autotune_over(h, iters, cand, tile_rows);
}

int sw2d_set_retune(sw2d_handle *h, long every, int iters)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (every < 0 || (every > 0 && iters < 1)) return fail(h, SW2D_E_ARG, "every >= 0 and iters >= 1");
    if (h->attached) return fail(h, SW2D_E_STATE, "attached batch members are tuned by sw2d_batch_autotune");
    h->retune_every = fused(h) ? every : 0;
    h->retune_iters = iters;
    h->next_retune = h->step_index + every;
    return SW2D_OK;
}

int sw2d_get_tile_rows(sw2d_handle *h, int *rows)
{
    if (check_idle(h) || !rows) return SW2D_E_ARG;
    *rows = fused(h) ? h->H : 0;
    return SW2D_OK;
}

int sw2d_set_tile_rows(sw2d_handle *h, int tile_rows)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (tile_rows < 1 || tile_rows > 512) return fail(h, SW2D_E_ARG, "tile_rows must be in [1, 512]");
    if (h->attached) return fail(h, SW2D_E_STATE, "attached batch members are tiled by their batch");
    if (!fused(h)) return SW2D_OK;
    return retile(h, tile_rows);
}

int sw2d_step(sw2d_handle *h, int n_iters)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (n_iters < 0) return fail(h, SW2D_E_ARG, "negative iteration count");
    if (h->attached) return fail(h, SW2D_E_STATE, "an attached batch member is stepped by sw2d_batch_step");
    if (!fused(h)) {
        for (int i = 0; i < n_iters; i++) {
            const StepCommon sc = make_common(h);
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (h->timing) {
                HIPCHK(h, hipEventCreate(&e0));
                HIPCHK(h, hipEventCreate(&e1));
                HIPCHK(h, hipEventRecord(e0, h->stream));
            }
            hipError_t e = launch_plain(h, sc);
            if (e != hipSuccess) return hip_fail(h, e, "step launch");
            if (h->timing) {
                HIPCHK(h, hipEventRecord(e1, h->stream));
                h->events.emplace_back(e0, e1);
                h->launches++;
                if (h->events.size() >= 1024 && flush_events(h)) return SW2D_E_HIP;
            }
            advance(h, 1);
        }
        return SW2D_OK;
    }
    int rc = upload_desc(h);
    if (rc) return rc;
    long left = n_iters;
    while (left > 0) {
      // periodic retuning: the current height against the round-filling ones
      if (h->retune_every > 0 && h->step_index >= h->next_retune) {
          std::vector<int> cand = fill_heights(h);
          cand.push_back(h->H);
          if ((rc = autotune_over(h, h->retune_iters, cand, nullptr))) return rc;
          if ((rc = upload_desc(h))) return rc;
          h->next_retune = h->step_index + h->retune_every;
      }
      long n = left;
      if (h->retune_every > 0) n = std::min<long>(n, std::max<long>(1, h->next_retune - h->step_index));
      left -= n;
      while (n > 0) {
        // whole groups of six from phase 0 replay the captured graph
        const bool group = h->graph_on && (h->step_index % 6 == 0) && n >= 6;
        if (group && (rc = ensure_graph(h))) return rc;
        const long m = group ? 6 * (n / 6) : 1;
        const long reps = group ? m / 6 : 1;
        for (long r = 0; r < reps; r++) {
            hipEvent_t e0 = nullptr, e1 = nullptr;
            if (h->timing) {
                HIPCHK(h, hipEventCreate(&e0));
                HIPCHK(h, hipEventCreate(&e1));
                HIPCHK(h, hipEventRecord(e0, h->stream));
            }
            if (group) {
                HIPCHK(h, hipGraphLaunch(h->gexec, h->stream));
            } else {
                hipError_t e = launch_one(h, h->dscen, (int)(h->step_index % 6));
                if (e != hipSuccess) return hip_fail(h, e, "step launch");
            }
            if (h->timing) {
                HIPCHK(h, hipEventRecord(e1, h->stream));
                h->events.emplace_back(e0, e1);
                h->launches += group ? 6 : 1;
                if (h->events.size() >= 1024 && flush_events(h)) return SW2D_E_HIP;
            }
        }
        advance(h, m);
        n -= m;
      }
    }
    return SW2D_OK;
}

int sw2d_sync(sw2d_handle *h)
{
    if (check_idle(h)) return SW2D_E_ARG;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return SW2D_OK;
}

int sw2d_time(sw2d_handle *h, double *time, long *mstep, double *dt2_next)
{
    if (check_idle(h)) return SW2D_E_ARG;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    Clock ck;
    HIPCHK(h, hipMemcpy(&ck, h->clk + (h->step_index % 3), sizeof(Clock), hipMemcpyDeviceToHost));
    if (time) *time = ck.t;
    if (mstep) *mstep = (long)h->host_mstep;
    if (dt2_next) {
        if (h->prm.adaptive) {
            const double m = __builtin_bit_cast(double, ck.cfl_bits);
            const double d = h->prm.courant * m;
            const double mx = 2.00 * h->prm.dt;
            *dt2_next = (d < mx) ? d : mx;
        } else {
            *dt2_next = 2.00 * h->prm.dt;
        }
    }
    return SW2D_OK;
}

int sw2d_run_until(sw2d_handle *h, double tmax, long max_iters, long *done)
{
    if (check_idle(h)) return SW2D_E_ARG;
    long n = 0;
    if (!h->prm.adaptive) {
        // the host mirror of the clock is exact in fixed-dt mode
        double t = h->host_time;
        long cnt = 0;
        while (cnt < max_iters && t + h->prm.dt <= tmax) {
            t = t + h->prm.dt;
            t = t + h->prm.dt;
            cnt++;
        }
        while (n < cnt) {
            const int chunk = (int)std::min<long>(cnt - n, 1 << 20);
            int rc = sw2d_step(h, chunk);
            if (rc) return rc;
            n += chunk;
        }
    } else {
        const double dt2_max = 2.00 * h->prm.dt;
        for (;;) {
            if (n >= max_iters) break;
            double t, dt2;
            int rc = sw2d_time(h, &t, nullptr, &dt2);
            if (rc) return rc;
            if (!(t + 0.5 * dt2 <= tmax)) break;
            long safe = (long)floor((tmax - t) / dt2_max);
            if (safe < 1) safe = 1;
            safe = std::min(safe, max_iters - n);
            safe = std::min<long>(safe, 1 << 16);
            if ((rc = sw2d_step(h, (int)safe))) return rc;
            n += safe;
        }
    }
    if (done) *done = n;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    return SW2D_OK;
}

int sw2d_get_field(sw2d_handle *h, int field, double *out)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (!out) return fail(h, SW2D_E_ARG, "out is NULL");
    if (field < 0 || field > SW2D_FIELD_VV1) return fail(h, SW2D_E_ARG, "unknown field");
    const size_t RC = (size_t)h->R * h->C;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    if (fused(h)) {
        // derived fields go through an R*C FP64 buffer made on first use
        int rc;
        if (!h->scratch && (rc = dalloc(h, &h->scratch, RC))) return rc;
        hipError_t e = launch_derive(h, field, h->scratch, 0, (long)RC);
        if (e != hipSuccess) return hip_fail(h, e, "derive");
        HIPCHK(h, hipMemcpyAsync(out, h->scratch, RC * 8, hipMemcpyDeviceToHost, h->stream));
        HIPCHK(h, hipStreamSynchronize(h->stream));
        if (field == SW2D_FIELD_Z)
            for (size_t i = 0; i < RC; i++)
                if (out[i] != out[i]) out[i] = h->grid.nodata;
        return SW2D_OK;
    }
    const double *src = nullptr;
    switch (field) {
    case SW2D_FIELD_H: src = h->ph; break;
    case SW2D_FIELD_UM: src = h->pum; break;
    case SW2D_FIELD_VN: src = h->pvn; break;
    case SW2D_FIELD_UU: src = h->puu; break;
    case SW2D_FIELD_VV: src = h->pvv; break;
    case SW2D_FIELD_UUA: src = h->puua; break;
    case SW2D_FIELD_VVA: src = h->pvva; break;
    case SW2D_FIELD_Z: src = h->pz; break;
    case SW2D_FIELD_HM: src = h->phm; break;
    case SW2D_FIELD_HN: src = h->phn; break;
    case SW2D_FIELD_UU1: src = h->puu1; break;
    case SW2D_FIELD_VV1: src = h->pvv1; break;
    }
    HIPCHK(h, hipMemcpy(out, src, RC * 8, hipMemcpyDeviceToHost));
    return SW2D_OK;
}

int sw2d_balance(sw2d_handle *h, sw2d_balance_out *out)
{
    if (check_idle(h)) return SW2D_E_ARG;
    if (!out) return fail(h, SW2D_E_ARG, "out is NULL");
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const StepCommon sc = make_common(h);
    hipError_t e = launch_peek(h, sc);
    if (e != hipSuccess) return hip_fail(h, e, "ledger peek");
    Ledger L;
    HIPCHK(h, hipMemcpyAsync(&L, h->peek, sizeof(L), hipMemcpyDeviceToHost, h->stream));
    HIPCHK(h, hipStreamSynchronize(h->stream));
    out->volume = (h->step_index > 0) ? L.vol_last : h->volume0;
    out->volume0 = h->volume0;
    out->v_source = L.v_src;
    out->v_clip = L.v_clip;
    out->v_outflow = L.v_out;
    out->q_out = L.q_last;
    out->q_peak = L.q_peak;
    out->t_peak = L.t_peak;
    out->error = out->volume - (out->volume0 + out->v_source + out->v_clip - out->v_outflow);
    return SW2D_OK;
}

int sw2d_hydrograph(sw2d_handle *h, double *out, long cap, long *n)
{
    if (check_idle(h)) return SW2D_E_ARG;
    sw2d_balance_out b;
    int rc = sw2d_balance(h, &b);  // writes the pending sample of the last iteration
    if (rc) return rc;
    long have = (long)std::min<long long>(h->step_index, h->hydro_cap);
    long m = std::min(have, cap);
    if (m > 0 && out) HIPCHK(h, hipMemcpy(out, h->hydro, sizeof(double) * 2 * m, hipMemcpyDeviceToHost));
    if (n) *n = have;
    return SW2D_OK;
}

int sw2d_lake_inflow(sw2d_handle *h, double *out, int cap, int *n)
{
    if (check_idle(h)) return SW2D_E_ARG;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    const int m = std::min(cap, h->nbc);
    if (m > 0 && out) HIPCHK(h, hipMemcpy(out, h->bc_acc, sizeof(double) * m, hipMemcpyDeviceToHost));
    if (n) *n = h->nbc;
    return SW2D_OK;
}

int sw2d_timing(sw2d_handle *h, int enable)
{
    if (check_idle(h)) return SW2D_E_ARG;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (auto &ev : h->events) {
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    h->events.clear();
    h->kernel_ms = 0.0;
    h->launches = 0;
    h->timing = enable ? 1 : 0;
    return SW2D_OK;
}

int sw2d_kernel_time(sw2d_handle *h, double *ms, long *launches)
{
    if (check_idle(h)) return SW2D_E_ARG;
    HIPCHK(h, hipStreamSynchronize(h->stream));
    for (auto &ev : h->events) {
        float m = 0;
        HIPCHK(h, hipEventElapsedTime(&m, ev.first, ev.second));
        h->kernel_ms += m;
        (void)hipEventDestroy(ev.first);
        (void)hipEventDestroy(ev.second);
    }
    h->events.clear();
    // graph replays are timed as a whole: ms / launches is then the mean
    // launch-to-launch time of the six kernels
    if (ms) *ms = h->kernel_ms;
    if (launches) *launches = h->launches;
    return SW2D_OK;
}

int sw2d_kernel_features(sw2d_handle *h, int *features, int *variant)
{
    if (check_idle(h)) return SW2D_E_ARG;
    const int f = features_of(make_common(h));
    if (features) *features = f;
    if (variant) *variant = fused(h) ? fused_variant(f) : 0;
    return SW2D_OK;
}

int sw2d_device_bytes(sw2d_handle *h, long long *bytes)
{
    if (check_idle(h) || !bytes) return SW2D_E_ARG;
    long long b = 0;
    const size_t RC = (size_t)h->R * h->C;
    const long long RP = (long long)h->R * h->P;
    if (fused(h))
        b += 2 * RP * (depth_bytes(h->prm.precision) + 2 * flux_bytes(h->prm.precision)) + RP * 8 +
             (h->dz ? RP * 8 : 0) + (h->scratch ? (long long)RC * 8 : 0) + (h->cls ? RP : 0);
    else b += (long long)16 * h->NA * 8;
    b += (long long)2 * h->nparts * sizeof(Partial);
    *bytes = b;
    return SW2D_OK;
}

}  // extern "C"

// ---- device math probe ----
namespace {
__global__ void math_probe(int fn, const double *x, const double *y, double *out, int n)
{
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    // every lane stays active (the cross-lane moves need the whole wave)
    const double a = (i < n) ? x[i] : 0.0;
    const double b = (y && i < n) ? y[i] : 0.0;
    double r = 0;
    switch (fn) {
    case 0: r = pow1333(a); break;
    case 1: r = exp_d(a); break;
    case 2: r = log_d(a); break;
    case 3: r = __builtin_sqrt(a); break;
    case 4: r = a / b; break;
    case 5: r = __builtin_fmax(a, b); break;
    case 6: r = from_west(a); break;  // lane layout of the cross-lane moves
    case 7: r = from_east(a); break;
    default: r = 0; break;
    }
    if (i < n) out[i] = r;
}
}  // namespace

extern "C" int sw2d_device_math(int device, int fn, const double *x, const double *y, double *out, int n)
{
    if (!x || !out || n < 0 || fn < 0 || fn > 7) return SW2D_E_ARG;
    if ((fn == 4 || fn == 5) && !y) return SW2D_E_ARG;
    if (n == 0) return SW2D_OK;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SW2D_E_NODEV;
    if (hipSetDevice(device) != hipSuccess) return SW2D_E_HIP;
    double *dx = nullptr, *dy = nullptr, *dout = nullptr;
    const size_t bytes = sizeof(double) * n;
    int rc = SW2D_OK;
    if (hipMalloc(&dx, bytes) != hipSuccess || hipMalloc(&dout, bytes) != hipSuccess ||
        (y && hipMalloc(&dy, bytes) != hipSuccess))
        rc = SW2D_E_NOMEM;
    if (!rc && hipMemcpy(dx, x, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = SW2D_E_HIP;
    if (!rc && y && hipMemcpy(dy, y, bytes, hipMemcpyHostToDevice) != hipSuccess) rc = SW2D_E_HIP;
    if (!rc) {
        hipLaunchKernelGGL(math_probe, dim3((n + 255) / 256), dim3(256), 0, 0, fn, dx, dy, dout, n);
        if (hipGetLastError() != hipSuccess || hipDeviceSynchronize() != hipSuccess) rc = SW2D_E_HIP;
    }
    if (!rc && hipMemcpy(out, dout, bytes, hipMemcpyDeviceToHost) != hipSuccess) rc = SW2D_E_HIP;
    if (dx) (void)hipFree(dx);
    if (dy) (void)hipFree(dy);
    if (dout) (void)hipFree(dout);
    return rc;
}

namespace {
// streaming device-to-device copy of n 16-byte elements: each thread moves
// four elements a block-width apart per step (four 16-byte loads in flight
// before the stores; stores bypass the caches), grid-stride over the rest
__global__ void __launch_bounds__(256) stream_copy(const double2 *__restrict__ in,
                                                   double2 *__restrict__ out, size_t n)
{
    const size_t stride = (size_t)gridDim.x * blockDim.x * 4;
    size_t i = (size_t)blockIdx.x * blockDim.x * 4 + threadIdx.x;
    for (; i + 3 * blockDim.x < n; i += stride) {
        const double2 a = in[i], b = in[i + blockDim.x], c = in[i + 2 * blockDim.x],
                      d = in[i + 3 * blockDim.x];
        __builtin_nontemporal_store(a.x, &out[i].x); __builtin_nontemporal_store(a.y, &out[i].y);
        __builtin_nontemporal_store(b.x, &out[i + blockDim.x].x);
        __builtin_nontemporal_store(b.y, &out[i + blockDim.x].y);
        __builtin_nontemporal_store(c.x, &out[i + 2 * blockDim.x].x);
        __builtin_nontemporal_store(c.y, &out[i + 2 * blockDim.x].y);
        __builtin_nontemporal_store(d.x, &out[i + 3 * blockDim.x].x);
        __builtin_nontemporal_store(d.y, &out[i + 3 * blockDim.x].y);
    }
    for (; i < n; i += blockDim.x) out[i] = in[i];
}
}  // namespace

extern "C" int sw2d_stream_copy(int device, size_t bytes, int reps, double *gbs)
{
    if (!gbs || reps < 1 || bytes < 4096) return SW2D_E_ARG;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || device < 0 || device >= ndev) return SW2D_E_NODEV;
    if (hipSetDevice(device) != hipSuccess) return SW2D_E_HIP;
    const size_t n = bytes / sizeof(double2);
    double2 *in = nullptr, *out = nullptr;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    int rc = SW2D_OK;
    if (hipMalloc(&in, n * sizeof(double2)) != hipSuccess ||
        hipMalloc(&out, n * sizeof(double2)) != hipSuccess)
        rc = SW2D_E_NOMEM;
    if (!rc && (hipMemset(in, 0, n * sizeof(double2)) != hipSuccess ||
                hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess))
        rc = SW2D_E_HIP;
    float ms = 0;
    if (!rc) {
        const unsigned blocks = (unsigned)std::min<size_t>((n + 1023) / 1024, 16384);
        hipLaunchKernelGGL(stream_copy, dim3(blocks), dim3(256), 0, 0, in, out, n);  // warm-up
        (void)hipEventRecord(e0, 0);
        for (int r = 0; r < reps; r++)
            hipLaunchKernelGGL(stream_copy, dim3(blocks), dim3(256), 0, 0, in, out, n);
        (void)hipEventRecord(e1, 0);
        if (hipGetLastError() != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0)
            rc = SW2D_E_HIP;
    }
    if (!rc) *gbs = 2.0 * (double)(n * sizeof(double2)) * reps / (ms * 1e-3) / 1e9;
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    if (in) (void)hipFree(in);
    if (out) (void)hipFree(out);
    return rc;
}
