// sw2d_server.hip -- continuous-batching request loop over one sw2d_batch
// (C ABI in include/sw2d.h).
//
// A device thread owns the batch.  Between slices of `slice_iters`
// iterations it (1) retires members whose done flag is set, copying their
// hydrograph and ledger to the request and returning their arrays to the
// pool, and (2) admits queued requests in arrival order while the pool can
// hold them.  Everything admitted is stepped by the same launches, so a
// request that arrives while others run starts at the next slice boundary
// instead of waiting for the GPU to drain.  With adaptive tiling the running
// set is retiled after each change of membership to the height that fills
// whole rounds of wave slots: a few small requests run as many short tiles
// (their time per iteration is the tile's row chain), a full GPU as fewer,
// taller ones (fewer halo rows re-read).
#include <math.h>
#include <stdlib.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "sw2d_ops.hpp"

using namespace sw2d;

namespace {

struct Request {
    long id = 0;
    sw2d_request req{};
    sw2d::HostImages img;        // bed and initial depths, built at submission
    std::vector<double> rain, mtab;
    std::vector<unsigned char> mcls;
    sw2d_result res{};
    std::vector<double> hydro;   // (t, q) pairs after completion
    sw2d_handle *h = nullptr;    // while running
    bool cancel = false;         // asked to stop; the device thread drops it
    size_t footprint = 0;
    long cap = 0;
    std::string err;
};

double percentile(std::vector<double> v, double p)
{
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    // nearest-rank percentile
    size_t k = (size_t)ceil(p / 100.0 * (double)v.size());
    if (k < 1) k = 1;
    if (k > v.size()) k = v.size();
    return v[k - 1];
}

}  // namespace

struct sw2d_server {
    int device = 0, precision = SW2D_FP32, tile_rows = 32, slice = 60;
    bool adaptive = false;     // tile height follows the running set
    bool refit = false;        // membership changed since the last fit
    int cur_tile = 32;
    long retiles = 0;
    long max_active = 0;       // cell budget of the running set (0 = none)
    long active_cells = 0;     // cells of the running scenarios
    sw2d_batch *batch = nullptr;
    std::thread thr;
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    bool stop = false;
    bool dead = false;         // a batch step failed: nothing more will run
    // fault injection for tests (SW2D_FAULT_SERVER_STEP=k in the environment
    // at create): slice k asks the batch for a negative iteration count,
    // which sw2d_batch_step rejects like any failed step (-1: off)
    long fault_slice = -1;
    long slices = 0;
    std::deque<Request *> queue;
    std::vector<Request *> running;
    std::map<long, std::unique_ptr<Request>> all;
    long next_id = 0;
    long submitted = 0, completed = 0, failed = 0, cancelled = 0;
    double cell_updates = 0.0, busy = 0.0;
    // device-thread snapshots for the stats (read by other threads under mu)
    size_t pool_peak = 0, pool_cap = 0;
    long launches = 0;
    double kernel_ms = 0.0;
    std::vector<double> lat, wait;
    std::chrono::steady_clock::time_point t0;
    std::string err;

    static long cells_of(const Request *r) { return (long)r->req.grid.rows * r->req.grid.cols; }

    double now() const
    {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    // device thread: create, configure and attach a member for r
    int admit(Request *r)
    {
        const sw2d_request &q = r->req;
        sw2d_params p = q.params;
        p.precision = precision;
        p.kernel = SW2D_KERNEL_FUSED;
        p.device = device;
        sw2d_handle *h = nullptr;
        int rc = sw2d::create_impl(&q.grid, nullptr, nullptr, &p, batch, &h, &r->img);
        if (rc) { r->err = sw2d::g_create_error; return rc; }
        r->img = sw2d::HostImages();  // on the device now
        auto bad = [&](int code) {
            r->err = sw2d_last_error(h);
            sw2d_destroy(h);
            return code;
        };
        if (!r->rain.empty() && (rc = sw2d_set_rain_series(h, r->rain.data(), (int)r->rain.size(), q.dtrain)))
            return bad(rc);
        if (q.outlet_row >= 0 && (rc = sw2d_set_outlet(h, q.outlet_row, q.outlet_col))) return bad(rc);
        if (!r->mcls.empty() &&
            (rc = sw2d_set_manning_map(h, r->mcls.data(), r->mtab.data(), (int)r->mtab.size())))
            return bad(rc);
        if ((rc = sw2d_set_hydrograph(h, r->cap))) return bad(rc);
        if ((rc = sw2d_batch_attach(batch, h, q.max_iters, q.t_end))) {
            r->err = sw2d_batch_last_error(batch);
            sw2d_destroy(h);
            return rc;
        }
        r->h = h;
        return SW2D_OK;
    }

    // device thread: a cancelled running request leaves the batch, its
    // arrays go back to the pool, no result is kept
    void drop(Request *r)
    {
        sw2d_destroy(r->h);
        refit = true;
        std::lock_guard<std::mutex> lk(mu);
        r->h = nullptr;
        active_cells -= cells_of(r);
        r->res.state = SW2D_REQ_CANCELLED;
        r->res.t_done = now();
        cancelled++;
    }

    void retire(Request *r)
    {
        sw2d_handle *h = r->h;
        int done = 0;
        long iters = 0;
        double t = 0;
        (void)sw2d_batch_status(batch, h, &done, &iters, &t);
        sw2d_balance_out bal{};
        int rc = sw2d_balance(h, &bal);
        long n = 0;
        std::vector<double> tq(2 * (size_t)std::max<long>(r->cap, 1));
        if (!rc) rc = sw2d_hydrograph(h, tq.data(), r->cap, &n);
        const double cells = (double)r->req.grid.rows * (double)r->req.grid.cols;
        std::string e = rc ? sw2d_last_error(h) : "";
        sw2d_destroy(h);
        refit = true;
        std::lock_guard<std::mutex> lk(mu);
        r->h = nullptr;
        active_cells -= cells_of(r);
        r->res.iterations = iters;
        r->res.time = t;
        r->res.t_done = now();
        if (rc) {
            r->res.state = SW2D_REQ_FAILED;
            r->err = e;
            failed++;
        } else {
            const long m = std::min(n, r->cap);
            tq.resize(2 * (size_t)m);
            r->hydro.swap(tq);
            r->res.hydro_n = m;
            r->res.balance = bal;
            r->res.state = SW2D_REQ_DONE;
            completed++;
            cell_updates += cells * (double)iters;
            lat.push_back(r->res.t_done - r->res.t_submit);
            wait.push_back(r->res.t_admit - r->res.t_submit);
        }
    }

    void loop()
    {
        (void)hipSetDevice(device);
        for (;;) {
            // admission, in arrival order, while the pool can hold the head
            for (;;) {
                Request *r = nullptr;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    if (stop) return;
                    if (queue.empty()) break;
                    r = queue.front();
                    size_t used = 0, cap = 0;
                    sw2d_batch_memory(batch, &used, nullptr, &cap);
                    if (!running.empty()) {
                        if (r->footprint > cap - used) break;  // HBM footprint
                        if (max_active > 0 && active_cells + cells_of(r) > max_active) break;
                    }
                    queue.pop_front();
                }
                const int rc = admit(r);
                std::lock_guard<std::mutex> lk(mu);
                if (rc == SW2D_E_NOMEM && !running.empty()) {
                    queue.push_front(r);  // fragmented pool: wait for a retirement
                    break;
                }
                if (rc) {
                    r->res.state = SW2D_REQ_FAILED;
                    r->res.t_done = now();
                    failed++;
                    cv_done.notify_all();
                    continue;
                }
                r->res.state = SW2D_REQ_RUNNING;
                r->res.t_admit = now();
                active_cells += cells_of(r);
                running.push_back(r);
                refit = true;
            }
            if (adaptive && refit && !running.empty()) {
                // a failed retile leaves that member at its height; a HIP
                // error shows again at the next step
                int H = cur_tile, n = 0;
                (void)batch_fit_tiles(batch, &H, &n);
                refit = false;
                std::lock_guard<std::mutex> lk(mu);
                cur_tile = H;
                retiles += n;
            }
            if (running.empty()) {
                std::unique_lock<std::mutex> lk(mu);
                cv_work.wait_for(lk, std::chrono::milliseconds(20),
                                 [&] { return stop || !queue.empty(); });
                if (stop) return;
                continue;
            }
            const double ts = now();
            int nrun = 0;
            const int rc = sw2d_batch_step(batch, slices++ == fault_slice ? -1 : slice, &nrun);
            {
                double ms = 0;
                long nl = 0;
                size_t pk = 0, cp = 0;
                (void)sw2d_batch_kernel_time(batch, &ms, &nl);
                (void)sw2d_batch_memory(batch, nullptr, &pk, &cp);
                std::lock_guard<std::mutex> lk(mu);
                busy += now() - ts;
                kernel_ms = ms;
                launches = nl;
                pool_peak = pk;
                pool_cap = cp;
            }
            if (rc) {
                // the batch cannot be trusted any more: every request still
                // running or queued fails with the error, and submit refuses
                // new ones; waiters are woken
                std::lock_guard<std::mutex> lk(mu);
                err = std::string("batch step failed: ") + sw2d_batch_last_error(batch);
                const double t = now();
                for (Request *r : running) {
                    r->res.state = SW2D_REQ_FAILED;
                    r->res.t_done = t;
                    r->err = err;
                    active_cells -= cells_of(r);
                    failed++;
                }
                running.clear();
                for (Request *r : queue) {
                    r->res.state = SW2D_REQ_FAILED;
                    r->res.t_done = t;
                    r->err = err;
                    failed++;
                }
                queue.clear();
                dead = true;
                cv_done.notify_all();
                return;
            }
            std::vector<Request *> keep;
            for (Request *r : running) {
                bool cancel;
                {
                    std::lock_guard<std::mutex> lk(mu);
                    cancel = r->cancel;
                }
                if (cancel) {
                    drop(r);
                    continue;
                }
                int done = 0;
                (void)sw2d_batch_status(batch, r->h, &done, nullptr, nullptr);
                if (done) retire(r);
                else keep.push_back(r);
            }
            std::lock_guard<std::mutex> lk(mu);
            running.swap(keep);
            cv_done.notify_all();
        }
    }
};

extern "C" {

int sw2d_server_create(int device, int precision, size_t pool_bytes, int tile_rows, int slice_iters,
                       long max_active_cells, sw2d_server **out)
{
    if (!out) return SW2D_E_ARG;
    *out = nullptr;
    if (slice_iters < 0) {
        g_create_error = "slice_iters must not be negative";
        return SW2D_E_ARG;
    }
    sw2d_server *s = new sw2d_server();
    s->device = device;
    s->precision = precision;
    s->adaptive = tile_rows == 0;
    s->tile_rows = tile_rows > 0 ? tile_rows : 32;
    s->cur_tile = s->tile_rows;
    s->slice = slice_iters > 0 ? slice_iters : 60;
    s->max_active = max_active_cells > 0 ? max_active_cells : 0;
    int rc = sw2d_batch_create(device, precision, pool_bytes, s->tile_rows, &s->batch);
    if (rc) {
        delete s;
        return rc;
    }
    (void)sw2d_batch_timing(s->batch, 1);
    // requests of many sizes share each launch: their tiles are interleaved
    // block by block, so every stretch of the launch mixes long and short
    // scenarios (1-2% faster per launch on the served mix than each
    // request's tiles in one run)
    rc = sw2d_batch_set_schedule(s->batch, 1, 0);
    if (rc) {
        sw2d_batch_destroy(s->batch);
        delete s;
        return rc;
    }
    if (const char *f = getenv("SW2D_FAULT_SERVER_STEP")) s->fault_slice = atol(f);
    s->t0 = std::chrono::steady_clock::now();
    s->thr = std::thread([s] { s->loop(); });
    *out = s;
    return SW2D_OK;
}

int sw2d_server_submit(sw2d_server *s, const sw2d_request *q, long *id)
{
    if (!s || !q) return SW2D_E_ARG;
    const int R = q->grid.rows, C = q->grid.cols;
    if (R < 2 || C < 2 || !q->z) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "request needs a grid of at least 2x2 and z";
        return SW2D_E_ARG;
    }
    if (!(q->t_end == q->t_end) || (q->nrain > 0 && !q->rain_mm) || (q->nclass > 0 && (!q->manning_class || !q->manning_table))) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "bad request arrays or t_end";
        return SW2D_E_ARG;
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->dead) return SW2D_E_STATE;   // err holds the step failure
    }
    // a request's latency counts from the call, its image building included
    const double t_call = s->now();
    auto r = std::make_unique<Request>();
    r->req = *q;
    const size_t RC = (size_t)R * C;
    // the device arrays' host images are built here, on the submitting
    // thread: admission then only copies them, and the device thread goes
    // back to stepping sooner
    try {
        if (q->h0) {
            sw2d::prepare_images(&q->grid, q->z, q->h0, s->precision, r->img);
        } else {
            const std::vector<double> zero(RC, 0.0);
            sw2d::prepare_images(&q->grid, q->z, zero.data(), s->precision, r->img);
        }
    } catch (const std::bad_alloc &) {
        std::lock_guard<std::mutex> lk(s->mu);
        s->err = "host memory for the request's images";
        return SW2D_E_NOMEM;
    }
    if (q->nrain > 0) r->rain.assign(q->rain_mm, q->rain_mm + q->nrain);
    if (q->nclass > 0) {
        r->mcls.assign(q->manning_class, q->manning_class + RC);
        r->mtab.assign(q->manning_table, q->manning_table + q->nclass);
    }
    // pointers now refer to the copies
    r->req.z = r->req.h0 = r->req.rain_mm = r->req.manning_table = nullptr;
    r->req.manning_class = nullptr;
    long cap = q->hydro_cap;
    if (cap <= 0) {
        const double it = q->t_end / (2.0 * q->params.dt);
        cap = q->max_iters >= 0 ? q->max_iters : (long)std::min(ceil(it) + 16.0, 1.0e8);
    }
    r->cap = std::max<long>(cap, 1);
    sw2d_params p = q->params;
    p.precision = s->precision;
    // adaptive tiling may take a member down to 2-row tiles
    r->footprint = sw2d_batch_footprint(&q->grid, &p, s->adaptive ? 2 : s->tile_rows, r->cap, q->nclass > 0);
    std::lock_guard<std::mutex> lk(s->mu);
    r->id = s->next_id++;
    r->res.state = SW2D_REQ_QUEUED;
    r->res.t_submit = t_call;
    if (id) *id = r->id;
    s->queue.push_back(r.get());
    s->all[r->id] = std::move(r);
    s->submitted++;
    s->cv_work.notify_all();
    return SW2D_OK;
}

int sw2d_server_poll(sw2d_server *s, long id, sw2d_result *out)
{
    if (!s || !out) return SW2D_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) {
        s->err = "unknown request id";
        return SW2D_E_ARG;
    }
    *out = it->second->res;
    return SW2D_OK;
}

int sw2d_server_hydrograph(sw2d_server *s, long id, double *tq, long cap, long *n)
{
    if (!s) return SW2D_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end() || it->second->res.state != SW2D_REQ_DONE) {
        s->err = "no finished request with this id";
        return SW2D_E_STATE;
    }
    const std::vector<double> &v = it->second->hydro;
    const long m = (long)v.size() / 2;
    if (n) *n = m;
    if (tq && cap > 0) std::copy(v.begin(), v.begin() + 2 * std::min(m, cap), tq);
    return SW2D_OK;
}

int sw2d_server_release(sw2d_server *s, long id)
{
    if (!s) return SW2D_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) return SW2D_OK;
    const int st = it->second->res.state;
    if (st != SW2D_REQ_DONE && st != SW2D_REQ_FAILED && st != SW2D_REQ_CANCELLED) {
        s->err = "request has not finished";
        return SW2D_E_STATE;
    }
    s->all.erase(it);
    return SW2D_OK;
}

int sw2d_server_cancel(sw2d_server *s, long id)
{
    if (!s) return SW2D_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->all.find(id);
    if (it == s->all.end()) {
        s->err = "unknown request id";
        return SW2D_E_ARG;
    }
    Request *r = it->second.get();
    const int st = r->res.state;
    if (st != SW2D_REQ_QUEUED && st != SW2D_REQ_RUNNING) return SW2D_OK;  // already finished
    auto q = std::find(s->queue.begin(), s->queue.end(), r);
    if (q != s->queue.end()) {
        // still queued: it never reaches the GPU
        s->queue.erase(q);
        r->res.state = SW2D_REQ_CANCELLED;
        r->res.t_done = s->now();
        s->cancelled++;
        s->cv_done.notify_all();
        return SW2D_OK;
    }
    // running, or being admitted right now: the device thread drops it after
    // the slice in progress
    r->cancel = true;
    return SW2D_OK;
}

int sw2d_server_wait(sw2d_server *s, long id, double timeout_s)
{
    if (!s) return SW2D_E_ARG;
    std::unique_lock<std::mutex> lk(s->mu);
    auto finished = [&] {
        if (id >= 0) {
            auto it = s->all.find(id);
            if (it == s->all.end()) return true;
            const int st = it->second->res.state;
            return st == SW2D_REQ_DONE || st == SW2D_REQ_FAILED || st == SW2D_REQ_CANCELLED;
        }
        // every submitted request has finished; the queue and the running set
        // are both empty for a moment while the device thread admits a request
        return s->completed + s->failed + s->cancelled == s->submitted;
    };
    const bool ok = s->cv_done.wait_for(lk, std::chrono::duration<double>(timeout_s), finished);
    if (!ok) {
        s->err = "timed out";
        return SW2D_E_STATE;
    }
    return SW2D_OK;
}

int sw2d_server_stats_get(sw2d_server *s, sw2d_server_stats *out)
{
    if (!s || !out) return SW2D_E_ARG;
    std::lock_guard<std::mutex> lk(s->mu);
    memset(out, 0, sizeof *out);
    out->submitted = s->submitted;
    out->completed = s->completed;
    out->failed = s->failed;
    out->cancelled = s->cancelled;
    out->running = (long)s->running.size();
    out->queued = (long)s->queue.size();
    out->cell_updates = s->cell_updates;
    out->busy_seconds = s->busy;
    out->uptime = s->now();
    out->p50_latency = percentile(s->lat, 50);
    out->p99_latency = percentile(s->lat, 99);
    out->max_latency = s->lat.empty() ? 0.0 : *std::max_element(s->lat.begin(), s->lat.end());
    out->p50_wait = percentile(s->wait, 50);
    out->p99_wait = percentile(s->wait, 99);
    out->pool_peak = s->pool_peak;
    out->pool_capacity = s->pool_cap;
    out->launches = s->launches;
    out->kernel_ms = s->kernel_ms;
    out->tile_rows = s->cur_tile;
    out->retiles = s->retiles;
    return SW2D_OK;
}

const char *sw2d_server_last_error(const sw2d_server *s) { return s ? s->err.c_str() : g_create_error.c_str(); }

int sw2d_server_destroy(sw2d_server *s)
{
    if (!s) return SW2D_OK;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        s->cv_work.notify_all();
    }
    if (s->thr.joinable()) s->thr.join();
    (void)sw2d_batch_destroy(s->batch);
    delete s;
    return SW2D_OK;
}

}  // extern "C"
