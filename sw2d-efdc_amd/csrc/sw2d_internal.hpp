// sw2d_internal.hpp -- handle layout and kernel argument records of libsw2d.
//
// Device-resident time keeping: every iteration reads its start time, its
// length and the CFL minimum from a three-slot clock ring and writes the next
// slot, so a sequence of iterations needs no host round trip (and the loop
// can be replayed from a HIP graph).  Iteration k reads slot k%3, writes the
// time of slot (k+1)%3, folds its CFL candidates into slot (k+1)%3 with an
// atomic minimum, and resets the CFL field of slot (k+2)%3.
//
// Volume ledger: each tile (wave) of iteration k writes one record of FP64
// partial sums into partial buffer k%2; iteration k+1 reduces buffer (k)%2 in
// a fixed order, so the ledger is deterministic and independent of how the
// waves were scheduled.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <array>
#include <string>
#include <vector>

#include <stdlib.h>

#include "../../include/sw2d.h"
#include "sw2d_pool.hpp"

namespace sw2d {

// SW2D_GRAPHS=0 in the environment when a handle or batch is created: its
// iterations are launched one by one instead of as replays of the captured
// six-launch graph (for tools that trace launches one at a time); the
// results are the same either way
inline bool graphs_enabled()
{
    const char *e = getenv("SW2D_GRAPHS");
    return !(e && e[0] == '0');
}

// pointer into global memory (address space 1 of the AMDGPU target)
template <typename T>
using gptr = __attribute__((address_space(1))) T *;

constexpr unsigned long long kInfBits = 0x7ff0000000000000ULL;
constexpr int kMaxClasses = 8;

struct Clock {
    double t;                    // time at the start of the iteration reading this slot
    double dt2_prev;             // length of the iteration that wrote this slot
    unsigned long long cfl_bits; // CFL minimum as the bits of a positive double
    long long mstep;             // reference half-step counter
};

struct Partial {
    double vol;   // sum of h over active output cells (new state)
    double src;   // sum of dt2 * (ql + rr*P - E) over active output cells
    double clip;  // sum of -h_raw where the new depth was clipped at 0
    double qout;  // outlet discharge of this iteration (m3/s), one tile only
};

struct Ledger {
    double v_src, v_clip, v_out;
    double q_last, q_peak, t_peak;
    double vol_last;       // volume of the state after the last reduced iteration
    long long reduced;     // iterations reduced so far
};

// Everything an iteration needs besides the state arrays; shared by the fused
// and the plain kernels.  Plain data only (passed by value).
struct StepCommon {
    int R, C;
    double dx, dy, g, rn, th;
    double P_lake, P_land, EV_WL_min;
    int fixed_dt;
    double dt_fixed, dt2_fixed, courant, dt2_max;
    Clock *clk;
    int slot_a, slot_b, slot_c;
    long long step_index;
    // forcing tables
    const double *rain; int nrain; double dtrain;
    const double *evapo; int nevapo; int evap_on;
    const double *Q; int nq; int nst; double dtoq;
    const int *src_r; const int *src_c;
    // extensions
    int storm_on;
    double st_peak, st_x0, st_y0, st_vx, st_vy, st_sigma, st_t0, st_t1;
    int ncls_on;
    double ntab[kMaxClasses];
    int out_r, out_c, out_edge;  // edge 0 none, 1 south, 2 north, 3 west, 4 east
    // ledger and bookkeeping
    Partial *part_out;           // written by this iteration
    const Partial *part_in;      // written by the previous iteration
    int nparts;
    Ledger *ledger;
    double *hydro; long long hydro_cap;
    int nbc; const int *bc_r; const int *bc_c; const int *bc_kind; double *bc_acc;
};

// The iteration-dependent part of a launch: clock slots, iteration index and
// the halves of the partial-sum buffer.  Kept apart from StepCommon so that
// the fused kernel can read StepCommon in place from its descriptor.
struct Iter {
    int slot_a, slot_b, slot_c;
    long long step_index;
    Partial *part_out;
    const Partial *part_in;
};

__host__ __device__ inline Iter iter_of(const StepCommon &s)
{
    return Iter{s.slot_a, s.slot_b, s.slot_c, s.step_index, s.part_out, s.part_in};
}

// Progress of a scenario stepped inside a batch, written by its bookkeeping
// wave after every iteration it runs.
struct Status {
    long long k;     // iterations completed
    double t;        // time after them
    int done;        // set by the first launch whose iteration would pass k_end / t_end
    int fin;         // set with the last iteration when that is already known (the
                     // iteration bound, or the time bound with a fixed step)
};

// Device descriptor of one scenario for the fused kernel.  Everything that
// changes from one iteration to the next is derived inside the kernel from the
// launch phase (iteration k mod 6, which fixes the clock slots k%3.. and the
// ping-pong buffer k%2) and from the clock ring itself (k = mstep / 2), so a
// fixed sequence of six launches can be captured once and replayed.
struct Scenario {
    StepCommon s;            // slot_*, step_index and part_in/out are filled per launch
    void *st[2][3];          // (h, um, vn) of ping-pong buffer 0 and 1; iteration k reads k%2
    const double *z;         // bed elevation (FP64 in both forms); NaN marks an inactive cell
    const double *dz;        // FP32 form: per cell its FP32 bed steps from the north and the
                             // west neighbour (two floats in 8 bytes, NaN for an inactive
                             // cell), which its fused kernel reads instead of z
    const uint8_t *cls;      // Manning class per cell (ncls_on)
    int P;                   // row pitch in elements
    int W;                   // output columns per tile (64 - 6)
    int H;                   // output rows per tile
    int nstrips, ntiles;     // tiles of this scenario
    int nblocks;             // tiles + bookkeeping waves
    Partial *part;           // two halves of nparts records
    int phase0;              // launch phase of this scenario's iteration 0 (mod 6)
    long long k_end;         // a batch member stops before iteration k_end ...
    double t_end;            // ... or before an iteration that would end past t_end + dt
    Status *status;          // batch members only; NULL for a stand-alone handle
};

// one block of a batched launch: scenario and block within it
struct TileRef {
    int sid, local;
};

// Features of a scenario that the fused kernel can compile out where no
// scenario of a launch uses them.
enum : int {
    kFeatStorm = SW2D_FEAT_STORM,        // moving storm (sw2d_set_storm)
    kFeatSources = SW2D_FEAT_SOURCES,    // point inflow series (stream_flow sources)
    kFeatAdaptive = SW2D_FEAT_ADAPTIVE,  // CFL time step
    kFeatMap = SW2D_FEAT_MAP,            // Manning class map
    kFeatAll = 15,
};
// the instantiated feature sets, smallest first
constexpr int kVariant0 = 0;
constexpr int kVariant1 = kFeatMap;
constexpr int kVariant2 = kFeatAdaptive | kFeatMap;
constexpr int kVariant3 = kFeatStorm | kFeatAdaptive | kFeatMap;
inline int fused_variant(int features)
{
    const int v[4] = {kVariant0, kVariant1, kVariant2, kVariant3};
    for (int i = 0; i < 4; i++)
        if ((features & ~v[i]) == 0) return v[i];
    return kFeatAll;
}
inline int features_of(const StepCommon &s)
{
    return (s.storm_on ? kFeatStorm : 0) | (s.nst > 0 ? kFeatSources : 0) |
           (s.fixed_dt ? 0 : kFeatAdaptive) | (s.ncls_on ? kFeatMap : 0);
}

struct SnapRing;   // asynchronous output copies (sw2d_output.hip)

}  // namespace sw2d

struct sw2d_batch;

// the opaque handle of the C ABI
struct sw2d_handle {
    int R = 0, C = 0, P = 0;
    sw2d_grid_desc grid{};
    sw2d_params prm{};
    int device = 0;
    hipStream_t stream = nullptr;
    int features = sw2d::kFeatAll;      // kFeat* of the current descriptor
    long retune_every = 0;              // sw2d_set_retune: iterations between retunings (0 = off)
    int retune_iters = 3;
    long long next_retune = 0;
    // fused-path state (ping-pong)
    void *h[2] = {nullptr, nullptr}, *um[2] = {nullptr, nullptr}, *vn[2] = {nullptr, nullptr};
    void *z = nullptr;
    void *dz = nullptr;                 // FP32 form: bed steps (R x P pairs of floats)
    uint8_t *cls = nullptr;
    int cur = 0;
    int W = 58, H = 64, nstrips = 0, nchunks = 0, ntiles = 0;
    double *scratch = nullptr;          // R*P doubles for derived fields
    // plain-path state (the reference's flat layout, N = R*C + C/2 plus a zero tail)
    int N = 0, NA = 0, front = 0;
    double *pz = nullptr, *ph = nullptr, *pho = nullptr, *phm = nullptr, *phn = nullptr,
           *pum = nullptr, *pumo = nullptr, *puu1 = nullptr, *pvv1 = nullptr, *puua = nullptr,
           *pvva = nullptr, *pvn = nullptr, *pvno = nullptr, *puu = nullptr, *pvv = nullptr;
    int *pinf = nullptr;
    double *pncell = nullptr;           // per-cell Manning n (plain path)
    int plain_blocks = 0;
    // shared device data
    sw2d::Clock *clk = nullptr;
    sw2d::Partial *part = nullptr;      // 2 * nparts
    int nparts = 0;
    sw2d::Ledger *ledger = nullptr;
    sw2d::Ledger *peek = nullptr;       // ledger + pending partials, for the host
    double *rain = nullptr; int nrain = 0; double dtrain = 1.0;
    double *evapo = nullptr; int nevapo = 0; int evap_on = 0;
    double *Q = nullptr; int nq = 0, nst = 0; double dtoq = 1.0;
    int *src_r = nullptr, *src_c = nullptr;
    int storm_on = 0; sw2d_storm storm{};
    int ncls_on = 0; double ntab[sw2d::kMaxClasses] = {0};
    int out_r = -1, out_c = -1, out_edge = 0;
    double *hydro = nullptr; long long hydro_cap = 0;
    int nbc = 0; int *bc_r = nullptr, *bc_c = nullptr, *bc_kind = nullptr;
    double *bc_acc = nullptr;
    double volume0 = 0.0;
    // host mirror of the clock (exact in fixed-dt mode)
    long long step_index = 0;
    double host_time = 0.0;
    long long host_mstep = 0;
    // fused-path launch descriptor (device copy of make_scenario) and the
    // six-launch graph that replays iterations k..k+5 for k % 6 == 0
    sw2d::Scenario *dscen = nullptr;
    bool desc_dirty = true;
    bool graph_on = true;
    hipGraphExec_t gexec = nullptr;
    // batch membership (sw2d_batch_new): memory from the batch pool, the
    // batch's stream, stepped by sw2d_batch_step while attached
    sw2d_batch *batch = nullptr;
    bool attached = false;
    int member_slot = -1;
    int phase0 = 0;                     // batch launch phase of this member's iteration 0 (mod 6)
    long long k_end = 0;                // stop before this iteration ...
    double t_end = 0.0;                 // ... or before one that would end past t_end + dt
    // timing
    int timing = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    double kernel_ms = 0.0;
    long launches = 0;
    std::vector<void *> allocs;
    sw2d::SnapRing *snap = nullptr;     // sw2d_snapshot slots, made on first use
    std::string err;
};

// a batch: scenarios stepped together, one launch per iteration for all
struct sw2d_batch {
    int device = 0;
    int precision = SW2D_FP32;
    int features = 0;   // kFeat* over the attached members (OR)
    hipStream_t stream = nullptr;
    char *pool_base = nullptr;
    sw2d::Pool pool;
    int H = 0;                              // member tile rows (0: each member's default)
    // launch schedule (sw2d_batch_set_schedule): block order 0 = each
    // member's tiles in a run, 1 = members interleaved tile by tile; a cap
    // on resident waves per SIMD (0 = none), enforced by reserving lds_pad
    // bytes of unused LDS per one-wave block
    int order = 0;
    int waves_cap = 0;
    size_t lds_pad = 0;
    // trials of the last sw2d_batch_autotune: (tile rows, order, waves cap, ms)
    std::vector<std::array<double, 4>> tune_log;
    std::vector<sw2d_handle *> members;     // attached, in attach order
    std::vector<sw2d_handle *> owned;       // every handle made by sw2d_batch_new and not destroyed
    std::vector<int> free_slots;            // status / descriptor slots
    sw2d::Scenario *dscen = nullptr;        // kMaxMembers descriptors
    sw2d::Status *dstatus = nullptr;        // kMaxMembers status records
    sw2d::Status *hstatus = nullptr;        // pinned host copy
    sw2d::TileRef *dtiles = nullptr;
    int tiles_cap = 0;
    int nblocks = 0;
    bool dirty = true;                      // descriptor / tile tables need a rebuild
    long long phase = 0;                    // launches issued (mod 6 is the launch phase)
    hipGraphExec_t gexec = nullptr;
    bool graph_on = true;
    int timing = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    long long launches = 0;
    double kernel_ms = 0.0;
    std::string err;
};

namespace sw2d {
constexpr int kMaxMembers = 4096;
// launchers (sw2d_fused64.hip, sw2d_fused32.hip, sw2d_plain.hip)
// one iteration of every scenario in scen[] (tiles == NULL: the single
// scenario scen[0], blocks in order) at launch phase `phase`
// `features` (kFeat*) covers every scenario of the launch; the smallest
// instantiation of the kernel that handles them runs
// lds: dynamic LDS bytes per block reserved and left unused (a cap on the
// resident blocks per CU; 0 = none)
hipError_t launch_fused_f64(hipStream_t stream, const Scenario *scen, const TileRef *tiles, int nblocks,
                           int phase, int features, size_t lds = 0);
hipError_t launch_fused_f32(hipStream_t stream, const Scenario *scen, const TileRef *tiles, int nblocks,
                           int phase, int features, size_t lds = 0);
hipError_t launch_fused_mix(hipStream_t stream, const Scenario *scen, const TileRef *tiles, int nblocks,
                            int phase, int features, size_t lds = 0);
hipError_t launch_init_f64(sw2d_handle *h);
hipError_t launch_init_f32(sw2d_handle *h);
hipError_t launch_bed_steps(sw2d_handle *h);
hipError_t launch_init_mix(sw2d_handle *h);
// one derived field of cells first .. first + count - 1 (row-major, rows x
// cols) into out_dev[0 .. count - 1]
hipError_t launch_derive_f64(sw2d_handle *h, int field, double *out_dev, long first, long count);
hipError_t launch_derive_f32(sw2d_handle *h, int field, double *out_dev, long first, long count);
hipError_t launch_derive_mix(sw2d_handle *h, int field, double *out_dev, long first, long count);
hipError_t launch_peek(sw2d_handle *h, const StepCommon &sc);
int fused_waves_per_cu_f64(int features);  // resident one-wave blocks per CU
int fused_waves_per_cu_f32(int features);
int fused_waves_per_cu_mix(int features);

// bytes of a stored depth and of a stored face flux in a precision form
inline int depth_bytes(int prec) { return prec == SW2D_FP32 ? 4 : 8; }
inline int flux_bytes(int prec) { return prec == SW2D_FP64 ? 8 : 4; }
inline bool known_precision(int prec)
{
    return prec == SW2D_FP64 || prec == SW2D_FP32 || prec == SW2D_MIXED;
}
// the fused kernel of a precision form
inline hipError_t launch_fused(int prec, hipStream_t stream, const Scenario *scen,
                               const TileRef *tiles, int nblocks, int phase, int features,
                               size_t lds = 0)
{
    if (prec == SW2D_FP64) return launch_fused_f64(stream, scen, tiles, nblocks, phase, features, lds);
    if (prec == SW2D_FP32) return launch_fused_f32(stream, scen, tiles, nblocks, phase, features, lds);
    return launch_fused_mix(stream, scen, tiles, nblocks, phase, features, lds);
}
inline int fused_waves_per_cu(int prec, int features)
{
    if (prec == SW2D_FP64) return fused_waves_per_cu_f64(features);
    if (prec == SW2D_FP32) return fused_waves_per_cu_f32(features);
    return fused_waves_per_cu_mix(features);
}
hipError_t launch_plain(sw2d_handle *h, const StepCommon &sc);
hipError_t launch_plain_init(sw2d_handle *h);
}  // namespace sw2d
