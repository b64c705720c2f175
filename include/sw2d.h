/*
 * sw2d.h -- C ABI of the MI355X SW2D library (libsw2d.so).
 *
 * The reference (LabHidro/SW2D-EFDC, SW2D-EFDC.cu) has no library API: its
 * SW2D time step lives inside main() and is driven through raw device
 * pointers.  Each entry point below replaces one piece of that in-process
 * call sequence; the comment on each names the reference code it replaces.
 * Plain C: opaque handle, plain pointers and sizes, int status (0 = OK,
 * negative = error, text from sw2d_last_error), no exceptions cross it.
 *
 * Grid convention (the reference's): rows*cols cells in row-major memory
 * order, row 0 the north row of the ESRI grid, "um" the flux through the
 * north face of a cell (positive southwards, row direction, spacing dx),
 * "vn" the flux through its west face (positive eastwards, spacing dy).
 * Every field array passed in or out holds rows*cols doubles.
 *
 * Threading: one host thread per handle.  Work is queued on the handle's own
 * HIP stream; functions that return data synchronise that stream.
 */
#ifndef SW2D_H
#define SW2D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define SW2D_ABI_VERSION 1

/* status codes */
#define SW2D_OK 0
#define SW2D_E_ARG (-1)     /* bad argument */
#define SW2D_E_HIP (-2)     /* HIP runtime error */
#define SW2D_E_NOMEM (-3)   /* device (or host) allocation failed */
#define SW2D_E_STATE (-4)   /* call not valid in this state */
#define SW2D_E_IO (-5)      /* file missing or malformed */
#define SW2D_E_NODEV (-6)   /* no usable GPU */
#define SW2D_E_UNSUPPORTED (-7)

/* arithmetic of the state and fluxes */
#define SW2D_FP64 0   /* FP64 everywhere (parity mode) */
#define SW2D_FP32 1   /* FP32 state/fluxes, FP64 dt, volume and hydrograph */
#define SW2D_MIXED 2  /* FP32 fluxes and momentum arithmetic over FP64 depths: FP64 depth
                         storage, continuity and depth switches, dt, volume, hydrograph */

/* kernel organisation */
#define SW2D_KERNEL_FUSED 0  /* one fused, row-streaming kernel per iteration */
#define SW2D_KERNEL_PLAIN 1  /* the reference's unfused kernel sequence */

/* fields for sw2d_get_field */
#define SW2D_FIELD_H 0     /* depth h */
#define SW2D_FIELD_UM 1    /* row-direction face flux um */
#define SW2D_FIELD_VN 2    /* column-direction face flux vn */
#define SW2D_FIELD_UU 3    /* cell-centre velocity uu (VTK y_velocity = -uu) */
#define SW2D_FIELD_VV 4    /* cell-centre velocity vv (VTK x_velocity) */
#define SW2D_FIELD_UUA 5   /* uu1 averaged onto column faces */
#define SW2D_FIELD_VVA 6   /* vv1 averaged onto row faces */
#define SW2D_FIELD_Z 7     /* bed elevation as uploaded */
#define SW2D_FIELD_HM 8    /* row-face depth */
#define SW2D_FIELD_HN 9    /* column-face depth */
#define SW2D_FIELD_UU1 10  /* row-face velocity */
#define SW2D_FIELD_VV1 11  /* column-face velocity */

typedef struct sw2d_handle sw2d_handle;

/* dem.dat header (SW2D-EFDC.cu:3990-3995) */
typedef struct sw2d_grid_desc {
    int rows, cols;
    double xllcorner, yllcorner, cellsize, nodata;
} sw2d_grid_desc;

/* scalars of setup.dat and the constants main() hard-codes */
typedef struct sw2d_params {
    double g;          /* gravity, 9.80 (SW2D-EFDC.cu:3977) */
    double manning;    /* Manning n (setup.dat, 3789) */
    double th;         /* dry threshold, 1.0e-4 m (4644) */
    double dt;         /* TIME STEP, half an iteration (3771); dt2 = 2 dt (3976) */
    double time0;      /* start time in s (4230) */
    double INT, INF, LWL, EV_WL_min;  /* loss fractions, 3781-3787 */
    int precision;     /* SW2D_FP64 or SW2D_FP32 */
    int kernel;        /* SW2D_KERNEL_FUSED or SW2D_KERNEL_PLAIN */
    int adaptive;      /* extension: CFL-adaptive iteration length (0 = fixed dt) */
    double courant;    /* Courant number of the adaptive step (0.5) */
    int device;        /* HIP device ordinal */
    int tile_rows;     /* rows per tile of the fused kernel (0 = automatic) */
} sw2d_params;

/* moving Gaussian storm (extension): rate(x, y, t) = peak * gx * gy with
 * gx = exp(-(x - (x0 + vx t))^2 / (2 sigma^2)), gy likewise, for
 * t0 <= t < t1; x = (col + 0.5) cellsize, y = (row + 0.5) cellsize measured
 * from the north-west corner; peak in m/s.  Replaces the rain series. */
typedef struct sw2d_storm {
    double peak, x0, y0, vx, vy, sigma, t0, t1;
} sw2d_storm;

/* volume ledger (extension), all volumes in m3 */
typedef struct sw2d_balance_out {
    double volume;       /* sum of h * dx * dy over active cells now */
    double volume0;      /* same at creation */
    double v_source;     /* rain + point sources - evaporation admitted */
    double v_clip;       /* volume added by clipping negative depths to 0 */
    double v_outflow;    /* volume through the outlet weir */
    double q_out;        /* outlet discharge of the last iteration, m3/s */
    double q_peak;       /* peak outlet discharge, m3/s */
    double t_peak;       /* time of the peak, s */
    double error;        /* volume - (volume0 + v_source + v_clip - v_outflow) */
} sw2d_balance_out;

/* Fill params with the reference's defaults (g 9.80, th 1e-4, fused FP64). */
void sw2d_default_params(sw2d_params *p);

/* Library/ABI version (SW2D_ABI_VERSION). */
int sw2d_version(void);

/* Text of the last error on h, or of the last failed sw2d_create if h is NULL. */
const char *sw2d_last_error(const sw2d_handle *h);

/* Copy bed z and initial depth h0 to the device and apply the reference's
 * initialisation: init_inf + initiald (SW2D-EFDC.cu:29-91, launched at
 * 4763-4766) after the uploads of 4677-4729.  Cells with z == nodata are
 * inactive.  The library owns all device memory of the handle.  Grids of
 * 2^29 or more cells (cols padded to a multiple of 16) are refused with
 * SW2D_E_ARG: offsets on the device are 32-bit (byte offsets of doubles). */
int sw2d_create(const sw2d_grid_desc *grid, const double *z, const double *h0,
                const sw2d_params *params, sw2d_handle **out);

/* Free the handle and its device memory (replaces cudaFree, 5126-5160). */
int sw2d_destroy(sw2d_handle *h);

/* Rain series in mm per interval dtrain (rain.dat, 4078-4108).  Evaluated on
 * the device every half step with stream_flow's interpolation (546-553),
 * replacing the host call + H2D copy of rr at 4820-4822 and 4843-4846. */
int sw2d_set_rain_series(sw2d_handle *h, const double *rain_mm, int n, double dtrain);

/* Point sources: nst coordinates in metres (coord_source_sink.dat,
 * 4016-4023) and the Q table of nq rows x nst values in m3/s, spacing dtoq
 * (Q_source_sink.dat, 4041-4069).  The cell index uses stream_flow's formula
 * (541-544); values are evaluated on the device every half step (531-540),
 * replacing the dense ql upload of 4821 and 4845.  A source outside the
 * rows*cols real cells is an error. */
int sw2d_set_sources(sw2d_handle *h, const double *x, const double *y, int nst,
                     const double *Q, int nq, double dtoq);

/* Evaporation series in mm per interval dtrain (the output of
 * gpu_evaporation_calc, 556-586); enables E in continuity (626-633) through
 * evaporation_load (588-600), replacing the H2D of Ev at 4824-4827. */
int sw2d_set_evaporation(sw2d_handle *h, const double *evapo_mm, int n);

/* evaporation_calc semantics on the host: fills evapo_mm[n] from albedo,
 * temperatures T and solar radiation Rg (556-586). */
int sw2d_evaporation_series(double albedo, const double *T, const double *Rg, int n,
                            double dtime, double *evapo_mm);

/* Extension: per-cell Manning class (rows*cols bytes) and its table of nclass
 * (<= 8) coefficients; a face uses 0.5 * (n_west + n_east). */
int sw2d_set_manning_map(sw2d_handle *h, const unsigned char *cls, const double *table,
                         int nclass);

/* Extension: moving Gaussian storm (replaces the rain series). */
int sw2d_set_storm(sw2d_handle *h, const sw2d_storm *storm);

/* Extension: free-outflow outlet at an edge cell; its edge face passes the
 * weir discharge 0.5440 h sqrt(g h) of SW2D-EFDC.cu:131 (depth at the start
 * of the iteration), zeroed like any face when the cell dries. */
int sw2d_set_outlet(sw2d_handle *h, int row, int col);

/* Extension: record the outlet hydrograph, up to cap iterations. */
int sw2d_set_hydrograph(sw2d_handle *h, long cap);

/* Extension: lake-inflow accumulation.  mask holds rows*cols ints (mask.dat,
 * lake = 1); the records are those of couple_BC_efdc (728-1067) and each
 * record's inflow rate (the volumeFlow_BC == 1 formulas, 764-821) is
 * integrated over every iteration on the device. */
int sw2d_set_lake_bc(sw2d_handle *h, const int *mask);

/* Extension (the project's autotuner): time the fused kernel for each
 * candidate tile height (rows per one-wave tile) over iters launches on the
 * current state and keep the fastest; *tile_rows receives it (0 for the
 * plain kernels).  The trial launches write only the spare state buffer and
 * scratch bookkeeping, so it may run at any point of a run (the cost of a
 * tile height depends on how much of the grid is wet); the state does not
 * depend on the tile height.  Synchronises. */
int sw2d_autotune(sw2d_handle *h, int iters, int *tile_rows);

/* Extension: periodic retuning.  With every > 0, sw2d_step re-times the
 * current tile height against the heights that fill whole rounds of the
 * device's wave slots, `iters` trial launches each, every `every`
 * iterations, and retiles to the fastest (the wet area, and with it the
 * best height, changes as a flood develops).  The trial launches run inside
 * sw2d_step, which then synchronises; every = 0 turns it off. */
int sw2d_set_retune(sw2d_handle *h, long every, int iters);

/* The fused kernel's current tile height (rows per tile; 0 on the plain path). */
int sw2d_get_tile_rows(sw2d_handle *h, int *rows);

/* Extension: set the fused kernel's rows per tile (1..512) at any point of a
 * run, as sw2d_autotune does with its choice.  No effect on the plain
 * kernels; the state does not depend on it. */
int sw2d_set_tile_rows(sw2d_handle *h, int tile_rows);

/* Advance n loop iterations of the reference's call sequence (4798-4877):
 * flux, continuity(t), treat_error, continuity(t + dt), derived fields.
 * Asynchronous: returns once the work is queued.  Groups of six iterations
 * run as replays of a captured HIP graph; with SW2D_GRAPHS=0 in the
 * environment when the handle (or batch) is created, every iteration is its
 * own launch instead (for tools that trace launches one at a time), with the
 * same results. */
int sw2d_step(sw2d_handle *h, int n_iters);

/* Run iterations while time + dt <= tmax (the loop condition of 4798), at
 * most max_iters; *done receives the number run.  Synchronises. */
int sw2d_run_until(sw2d_handle *h, double tmax, long max_iters, long *done);

/* Wait for queued work. */
int sw2d_sync(sw2d_handle *h);

/* Simulated time (s), the reference's mstep (2 per iteration) and the
 * length dt2 of the next iteration.  Synchronises. */
int sw2d_time(sw2d_handle *h, double *time, long *mstep, double *dt2_next);

/* Copy one field (rows*cols doubles) to caller memory; derived fields are
 * computed from the current state as hm_hn/uu1_vv1/uu_vv/uua_vva do
 * (362-522).  Replaces the D2H copies at output times (4900, 4976, 5026,
 * 5042, 5058).  Synchronises. */
int sw2d_get_field(sw2d_handle *h, int field, double *out);

/* sw2d_get_field without any device allocation.  No counterpart in the
 * reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The same values as sw2d_get_field, but a derived field of a
 * fused handle is staged through a fixed 16 MiB device array that each GPU
 * holds from the library's load until the process ends, in chunks of whole
 * rows, instead of the R*C buffer sw2d_get_field allocates on first use.
 * Grids with more than 2^21 columns are refused (SW2D_E_ARG).  Calls on one
 * GPU share its array: the library serialises them within a process (a call
 * holds the array until its copy has finished), so calls that overlap on one
 * GPU wait for each other.  Synchronises; restores the caller's current
 * device. */
int sw2d_get_field_noalloc(sw2d_handle *h, int field, double *out);

/* Asynchronous output copies (double-buffered; stand-alone handles).  Same
 * values as sw2d_get_field, but the blocking copies at output times (4880-4916,
 * 4976-5072) become a snapshot that overlaps the steps that follow:
 * sw2d_snapshot queues, in stream order, the fields of `fields` (bit
 * 1u << SW2D_FIELD_x) at the current step and their copy into pinned host
 * memory on a copy stream, and returns without waiting.  Two snapshots may be
 * held at once; a third waits up to timeout_s for a release, then fails with
 * SW2D_E_STATE.  sw2d_snapshot_field waits for the copy and gives rows*cols
 * doubles valid until sw2d_snapshot_release.  sw2d_snapshot_field and
 * sw2d_snapshot_release may be called from another thread (a writer) while
 * this one steps the handle, but not during sw2d_destroy; they return a code
 * without setting the handle's error text (which belongs to the stepping
 * thread). */
int sw2d_snapshot(sw2d_handle *h, unsigned fields, double timeout_s, long *ticket);
int sw2d_snapshot_field(sw2d_handle *h, long ticket, int field, const double **data);
int sw2d_snapshot_release(sw2d_handle *h, long ticket);

/* couple_BC_efdc (728-1067) on caller-supplied host fields: records of the
 * four edge lists (S, N, E, W in the reference's naming) with the EFDC
 * (I, J) and inflow (m3/s).  Arrays cells/I/J/q hold 4*cap entries, list k
 * at offset k*cap; counts[4] receives the list lengths.  h, vva, uua are
 * rows*cols doubles (the reference passes h_vva as h_vn and h_uua as h_um,
 * 4903/4909).  Host-only. */
int sw2d_couple_bc(int rows, int cols, double dx, double dy, const int *mask, const double *h,
                   const double *vva, const double *uua, int cap, int *cells, int *I, int *J,
                   double *q, int *counts);

/* lake_catchment_smooth (2766-2827) on a host DEM, in place.  Host-only. */
int sw2d_lake_catchment_smooth(int rows, int cols, const int *mask, double *z);

/* Volume ledger (extension).  Synchronises. */
int sw2d_balance(sw2d_handle *h, sw2d_balance_out *out);

/* Outlet hydrograph: up to cap (time, Q) pairs; *n receives the count. */
int sw2d_hydrograph(sw2d_handle *h, double *out, long cap, long *n);

/* Accumulated lake inflow volume per record (order of sw2d_set_lake_bc:
 * all S records, then N, E, W); *n receives the record count. */
int sw2d_lake_inflow(sw2d_handle *h, double *out, int cap, int *n);

/* Timing of the main step kernel with HIP events on the handle's stream:
 * enable (1) or disable (0); sw2d_kernel_time returns the summed kernel
 * milliseconds and launch count since enabling.  Synchronises. */
int sw2d_timing(sw2d_handle *h, int enable);
int sw2d_kernel_time(sw2d_handle *h, double *ms, long *launches);

/* Which fused-kernel instantiation steps this handle (no counterpart in the
 * reference: a diagnostic of this library).  *features = SW2D_FEAT_* bits the
 * handle's scenario uses; *variant = the feature set of the instantiation
 * that runs it, the smallest one covering *features (features outside it are
 * compiled out).  A batch member runs the instantiation covering every member
 * of its batch; this reports its own scenario's. */
#define SW2D_FEAT_STORM 1
#define SW2D_FEAT_SOURCES 2
#define SW2D_FEAT_ADAPTIVE 4
#define SW2D_FEAT_MAP 8
int sw2d_kernel_features(sw2d_handle *h, int *features, int *variant);

/* Device memory the handle holds, in bytes. */
int sw2d_device_bytes(sw2d_handle *h, long long *bytes);

/* Evaluate the library's device math on n inputs (diagnostics and tests):
 * fn 0 pow(x, 1.3333330), 1 exp, 2 log, 3 sqrt, 4 division x/y,
 * 5 fmax(x, y), 6 the value of element i-1 within each group of 64
 * (the kernels' west-neighbour move; 0 for the first), 7 of element i+1
 * (east-neighbour move; 0 for the last).  x, y, out hold n doubles (y may be
 * NULL where unused). */
int sw2d_device_math(int device, int fn, const double *x, const double *y, double *out, int n);

/* Measured bandwidth ceiling (a diagnostic with no counterpart in the
 * reference): a streaming device-to-device copy of `bytes` bytes, 16-byte
 * loads and stores, run `reps` times after one warm-up and timed with HIP
 * events; *gbs = (bytes read + bytes written) / time, in GB/s. */
int sw2d_stream_copy(int device, size_t bytes, int reps, double *gbs);

/* ------------------------------------------------------------------------
 * Batches: many scenarios of any grid sizes advanced together, one kernel
 * launch per iteration for all of them (groups of six iterations replay a
 * captured HIP graph).  No counterpart in the reference, which runs one
 * grid per process (SW2D-EFDC.cu:3682 main); each member keeps the
 * single-grid semantics of sw2d_step exactly, so its results do not depend
 * on what else is in the batch.
 * ---------------------------------------------------------------------- */
typedef struct sw2d_batch sw2d_batch;

/* A batch on `device` owns one stream and a device memory pool of
 * pool_bytes; every member's arrays come from the pool.  All members share
 * `precision`.  tile_rows: rows per tile of members (0 = 32). */
int sw2d_batch_create(int device, int precision, size_t pool_bytes, int tile_rows, sw2d_batch **out);
/* Destroys the members still in the batch, then the batch. */
int sw2d_batch_destroy(sw2d_batch *b);
const char *sw2d_batch_last_error(const sw2d_batch *b);

/* Pool bytes a member with this grid and these params takes, with a
 * hydrograph of hydro_cap records and (manning_map != 0) a Manning class
 * map; tile_rows as for sw2d_batch_create. */
size_t sw2d_batch_footprint(const sw2d_grid_desc *grid, const sw2d_params *params, int tile_rows,
                            long hydro_cap, int manning_map);

/* A handle whose arrays live in the batch's pool and whose work runs on the
 * batch's stream.  All sw2d_* calls on handles apply to it; SW2D_E_NOMEM
 * when the pool cannot hold it.  params->kernel must be fused and
 * params->precision the batch's. */
int sw2d_batch_new(sw2d_batch *b, const sw2d_grid_desc *grid, const double *z, const double *h0,
                   const sw2d_params *params, sw2d_handle **out);

/* Join the batch's step loop.  The member stops before iteration
 * (current + max_iters) (max_iters < 0: no bound) and before the first
 * iteration whose t + dt exceeds t_end (INFINITY: no bound), the loop
 * condition of SW2D-EFDC.cu:4789.  While attached, sw2d_step on it is
 * SW2D_E_STATE. */
int sw2d_batch_attach(sw2d_batch *b, sw2d_handle *h, long max_iters, double t_end);
int sw2d_batch_detach(sw2d_batch *b, sw2d_handle *h);

/* Up to n_iters iterations of every attached member that has not finished;
 * *running = members still running afterwards.  Synchronises. */
int sw2d_batch_step(sw2d_batch *b, int n_iters, int *running);

/* Progress of a member after the last sw2d_batch_step. */
int sw2d_batch_status(sw2d_batch *b, const sw2d_handle *h, int *done, long *iters, double *time);

/* Measured choice of the members' tile rows, at any point of the run: each
 * candidate runs `iters` trial launches of the next iteration of every
 * running member (at the live launch phase) that write only spare buffers
 * and scratch bookkeeping; the members are then retiled (pending per-tile
 * partial sums are folded, so the ledger is unchanged). */
int sw2d_batch_autotune(sw2d_batch *b, int iters, int *tile_rows);

/* Launch schedule of the batch's step launches, the tuning dimensions beside
 * the tile height: order 0 lays each member's blocks out in one run
 * (neighbouring tiles share an XCD's L2), order 1 interleaves the members
 * block by block (every stretch of a launch mixes scenarios, which evens out
 * the work where some are wet and others dry); waves_per_simd > 0 caps the
 * resident one-wave blocks per SIMD, and so the tiles and scenarios sharing
 * a CU, by reserving unused LDS per block (0: the occupancy the registers
 * allow).  Results do not depend on either.  sw2d_batch_autotune times both
 * at its chosen tile height (orders 0 and 1, caps of one and two waves
 * below the register limit) and keeps a change only when it is at least
 * 1 % faster. */
int sw2d_batch_set_schedule(sw2d_batch *b, int order, int waves_per_simd);
int sw2d_batch_schedule(sw2d_batch *b, int *order, int *waves_per_simd);
/* The trials of the last sw2d_batch_autotune, in the order run: rows of
 * (tile rows, order, waves_per_simd, ms per launch), up to cap rows; *n
 * receives the count. */
int sw2d_batch_tune_report(sw2d_batch *b, double *rows, int cap, int *n);

/* Pool use in bytes: now, the most so far, capacity. */
int sw2d_batch_memory(sw2d_batch *b, size_t *in_use, size_t *peak, size_t *capacity);

/* HIP-event timing of the step launches (as sw2d_timing / sw2d_kernel_time). */
int sw2d_batch_timing(sw2d_batch *b, int enable);
int sw2d_batch_kernel_time(sw2d_batch *b, double *ms, long *launches);

/* ------------------------------------------------------------------------
 * Serving: a request loop over one batch, run by a device thread the server
 * owns.  Requests queue in arrival order and are admitted while the batch's
 * memory pool can hold them (HBM-footprint admission); admitted scenarios
 * are stepped together with those already running (continuous batching) and
 * retire when they reach their end time, leaving their outlet hydrograph and
 * volume ledger with the server.  No counterpart in the reference.
 * ---------------------------------------------------------------------- */
typedef struct sw2d_server sw2d_server;

typedef struct sw2d_request {
    sw2d_grid_desc grid;
    const double *z;                 /* rows*cols bed elevations */
    const double *h0;                /* rows*cols initial depths; NULL = dry */
    sw2d_params params;              /* precision and kernel are the server's */
    const double *rain_mm;           /* rain series (mm per dtrain), nrain values */
    int nrain;
    double dtrain;
    int outlet_row, outlet_col;      /* free-outflow outlet; -1 = none */
    const unsigned char *manning_class;  /* optional per-cell class map ... */
    const double *manning_table;         /* ... and its nclass coefficients */
    int nclass;
    double t_end;                    /* simulated end time, s */
    long max_iters;                  /* iteration bound (< 0 = none) */
    long hydro_cap;                  /* hydrograph records kept (0 = enough for t_end at dt) */
} sw2d_request;

#define SW2D_REQ_QUEUED 0
#define SW2D_REQ_RUNNING 1
#define SW2D_REQ_DONE 2
#define SW2D_REQ_FAILED (-1)
#define SW2D_REQ_CANCELLED 3

typedef struct sw2d_result {
    int state;                       /* SW2D_REQ_* */
    double t_submit, t_admit, t_done;  /* seconds on the server's clock; t_submit is
                                        * the submit call's start */
    long iterations;
    double time;                     /* simulated time reached */
    long hydro_n;                    /* hydrograph records held */
    sw2d_balance_out balance;
} sw2d_result;

typedef struct sw2d_server_stats {
    long submitted, completed, failed, running, queued;
    double cell_updates;             /* rows*cols*iterations of completed requests */
    double busy_seconds;             /* device-thread time spent stepping */
    double uptime;                   /* seconds since create */
    double p50_latency, p99_latency, max_latency;  /* t_done - t_submit, completed */
    double p50_wait, p99_wait;       /* t_admit - t_submit, completed */
    size_t pool_peak, pool_capacity;
    long launches;
    double kernel_ms;                /* HIP-event time of the step launches */
    long cancelled;
    long tile_rows;                  /* tile height new members start at */
    long retiles;                    /* members moved to another height (adaptive tiling) */
} sw2d_server_stats;

/* tile_rows: tile height of every request that names none (0 = adaptive:
 * after each admission or retirement the running requests are retiled to the
 * height that fills whole rounds of wave slots, the fewest rounds with a
 * height of at most 80, else 32; results are those of any fixed height, the
 * volume ledger to rounding).
 * slice_iters: iterations per batch step between admissions (0 = 60).
 * The server's batch interleaves the running requests' tiles block by block
 * (sw2d_batch_set_schedule order 1).
 * max_active_cells: admission also stops while the running scenarios hold
 * this many cells (0 = no limit; a request is always admitted onto an idle
 * GPU).  With enough cells in flight to fill the GPU, admitting more only
 * shares it among more requests: the same throughput, later results. */
int sw2d_server_create(int device, int precision, size_t pool_bytes, int tile_rows, int slice_iters,
                       long max_active_cells, sw2d_server **out);
/* Copies the request's arrays, the bed and initial depths already in the
 * layout of the device arrays (on the calling thread, so that admission on the
 * server's device thread is only the upload); returns its id.  Once a batch
 * step has failed the server stops: every running and queued request fails
 * with the error (sw2d_server_last_error) and submit returns SW2D_E_STATE. */
int sw2d_server_submit(sw2d_server *s, const sw2d_request *req, long *id);
int sw2d_server_poll(sw2d_server *s, long id, sw2d_result *out);
/* (t, q) pairs of a finished request: 2*min(n, cap) doubles. */
int sw2d_server_hydrograph(sw2d_server *s, long id, double *tq, long cap, long *n);
/* Forget a finished (done, failed or cancelled) request. */
int sw2d_server_release(sw2d_server *s, long id);
/* Stop a request: a queued one never runs; a running one leaves the batch
 * after the slice in progress and its device memory returns to the pool.
 * Its state becomes SW2D_REQ_CANCELLED (no hydrograph).  A finished request
 * is left as it is. */
int sw2d_server_cancel(sw2d_server *s, long id);
/* Wait until request id (id < 0: every submitted request) has finished,
 * failed or been cancelled; SW2D_E_STATE on timeout. */
int sw2d_server_wait(sw2d_server *s, long id, double timeout_s);
int sw2d_server_stats_get(sw2d_server *s, sw2d_server_stats *out);
const char *sw2d_server_last_error(const sw2d_server *s);
/* Stops the device thread (queued requests are dropped) and frees everything.
 * No other thread may be inside a sw2d_server_* call on s (a waiter must
 * have returned first). */
int sw2d_server_destroy(sw2d_server *s);

#ifdef __cplusplus
}
#endif

#endif /* SW2D_H */
