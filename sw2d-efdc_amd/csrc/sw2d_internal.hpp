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

#include <string>
#include <vector>

#include "../../include/sw2d.h"

namespace sw2d {

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

template <typename T>
struct FusedArgs {
    const T *h_in; const T *um_in; const T *vn_in;
    T *h_out; T *um_out; T *vn_out;
    const double *z;         // bed elevation (FP64 in both forms); NaN marks an inactive cell
    const uint8_t *cls;      // Manning class per cell (ncls_on)
    int P;                   // row pitch in elements
    int W;                   // output columns per tile (64 - 6)
    int H;                   // output rows per tile
    int nstrips, ntiles;
    int xcd_remap;
    StepCommon s;
};

}  // namespace sw2d

// the opaque handle of the C ABI
struct sw2d_handle {
    int R = 0, C = 0, P = 0;
    sw2d_grid_desc grid{};
    sw2d_params prm{};
    int device = 0;
    hipStream_t stream = nullptr;
    int esize = 8;                      // bytes per state element
    // fused-path state (ping-pong)
    void *h[2] = {nullptr, nullptr}, *um[2] = {nullptr, nullptr}, *vn[2] = {nullptr, nullptr};
    void *z = nullptr;
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
    // timing
    int timing = 0;
    std::vector<std::pair<hipEvent_t, hipEvent_t>> events;
    double kernel_ms = 0.0;
    long launches = 0;
    std::vector<void *> allocs;
    std::string err;
};

namespace sw2d {
// launchers (sw2d_fused64.hip, sw2d_fused32.hip, sw2d_plain.hip)
hipError_t launch_fused_f64(sw2d_handle *h, const StepCommon &sc);
hipError_t launch_fused_f32(sw2d_handle *h, const StepCommon &sc);
hipError_t launch_init_f64(sw2d_handle *h);
hipError_t launch_init_f32(sw2d_handle *h);
hipError_t launch_derive_f64(sw2d_handle *h, int field, double *out_dev);
hipError_t launch_derive_f32(sw2d_handle *h, int field, double *out_dev);
hipError_t launch_peek(sw2d_handle *h, const StepCommon &sc);
int fused_waves_per_cu_f64(int adaptive);  // resident one-wave blocks per CU
int fused_waves_per_cu_f32(int adaptive);
hipError_t launch_plain(sw2d_handle *h, const StepCommon &sc);
hipError_t launch_plain_init(sw2d_handle *h);
}  // namespace sw2d
