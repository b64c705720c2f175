// sw2d_fused_kernel.hpp -- one fused kernel per SW2D loop iteration.
//
// The reference (SW2D-EFDC.cu:4798-4877) runs nine grid-stride kernels per
// iteration over ~20 FP64 arrays: flux, continuity, treat_error, continuity,
// hm_hn, uu1_vv1, uu_vv, uua_vva, forward.  The complete state between
// iterations is only (h, um, vn) (bed z is constant); every other array is a
// local function of it.  This kernel reads (h, um, vn, z) once and writes
// (h, um, vn) once per iteration: 56 B/cell in FP64; 32 B/cell in the FP32
// form, whose bed is not rounded to binary32 (it is input data: rounding a
// DEM to binary32 moves it by up to half an ulp of its relief, which alone
// shifts depths by ~1 mm on a 2.5 m-relief test DEM, while FP32 state costs
// 0.2 mm): it reads, in the same 8 bytes, the cell's bed steps to its north
// and west neighbours, the FP64 differences rounded once at creation.
//
// Work decomposition.  A tile is one wavefront (64 lanes) and covers 58
// output columns x H output rows.  Lane l holds column j = j0 - 3 + l: the
// dependency chain flux(+-1) -> continuity (+-1) -> treat_error (+-1) ->
// continuity needs 3 cells of halo on each side.  The wave streams the rows
// of its tile from north to south; when row k arrives it computes
//   the face depths/velocities of row k            (hm_hn, uu1_vv1, uua_vva)
//   the fluxes of row k-1                          (flux)
//   the first continuity of row k-2                (continuity at t)
//   treat_error on the faces of row k-2
//   the second continuity of row k-3              (continuity at t + dt)
// and stores row k-3 of the new state.  Values of a row are carried in
// registers while they are needed (7-row window); east/west neighbours are
// DPP lane moves, made afresh in the row that uses them rather than carried
// to the next one, and the face depths of row k-1 are derived where they are
// used: register pressure sets the waves per SIMD.  No LDS staging and no
// barriers: each wave is independent.  The grid is one wave per tile plus the
// bookkeeping waves.
//
// T is the type of the fluxes and of the momentum arithmetic, TH that of the
// depths, the continuity updates and every switch that tests a depth.
// Instantiated three times: sw2d_fused64.hip (T = TH = double, built with
// -ffp-contract=off: every FP64 expression is written in the reference's
// evaluation order, so the results equal the reference arithmetic as restated
// by the CPU FP64 check, bit for bit), sw2d_fused32.hip (T = TH = float) and
// sw2d_fusedmix.hip (T = float, TH = double: FP32 fluxes and momentum
// arithmetic over FP64 depths, so the dry-bed and threshold switches see the
// depths the FP64 form sees).
#pragma once
#include "sw2d_common.hpp"

namespace sw2d {

// bit of lane-1 of a wave-wide predicate
__device__ __forceinline__ bool west_bit(bool p, int lane)
{
    unsigned long long m = __ballot(p);
    return lane > 0 && ((m >> (lane - 1)) & 1ULL);
}

// One scenario array, read and written at 32-bit element offsets.  Buf =
// true: through a buffer resource (base and range in scalar registers, a
// 32-bit byte offset per lane: no per-lane 64-bit address arithmetic; cell
// offsets stay below 2^29, which the host enforces, so byte offsets of
// doubles fit).  Buf = false: a global-address-space pointer (SGPR base +
// 64-bit VGPR offset).  Measured: buffers make the FP32 form's launches
// 1-2% faster and the FP64 and mixed forms' 1-2% slower, so only the FP32
// form uses them.
typedef unsigned int u32x2 __attribute__((ext_vector_type(2)));
template <typename V, bool Buf> struct Arr;
template <typename V> struct Arr<V, false> {
    gptr<V> p;
    __device__ explicit Arr(const void *q) : p((gptr<V>)const_cast<void *>(q)) {}
    __device__ __forceinline__ V load(unsigned off) const { return p[off]; }
    __device__ __forceinline__ void store(unsigned off, V v) const { p[off] = v; }
};
template <typename V> struct Arr<V, true> {
    __amdgpu_buffer_rsrc_t r;
    // range: the whole 32-bit offset space; 0x00020000: the 32-bit data
    // format (a buffer with the invalid format reads zeros)
    __device__ explicit Arr(const void *q)
        : r(__builtin_amdgcn_make_buffer_rsrc(const_cast<void *>(q), (short)0, -1, 0x00020000)) {}
    __device__ __forceinline__ V load(unsigned off) const
    {
        if constexpr (sizeof(V) == 4)
            return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b32(r, (int)(off * 4u), 0, 0));
        else
            return __builtin_bit_cast(V, __builtin_amdgcn_raw_buffer_load_b64(r, (int)(off * 8u), 0, 0));
    }
    __device__ __forceinline__ void store(unsigned off, V v) const
    {
        if constexpr (sizeof(V) == 4)
            __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(unsigned, v), r, (int)(off * 4u), 0, 0);
        else
            __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u32x2, v), r, (int)(off * 8u), 0, 0);
    }
};

constexpr int kMaxTileRows = 512;

// F: the features (kFeat*) this instantiation handles.  A feature missing
// from F is compiled out (its uniforms, branches and loops), which the host
// may only do for launches where no scenario uses it.  The arithmetic of
// what remains is the same in every instantiation (FP64 fuses no
// multiply-add, FP32 only within one source expression), so a scenario's
// results do not depend on the instantiation that ran it.
template <typename T, typename TH, int F>
__global__ void __launch_bounds__(64)
fused_step(const Scenario *__restrict__ scen, const TileRef *__restrict__ tiles, const int nblocks,
           const int phase)
{
    using M = num<T>;
    using MH = num<TH>;
    constexpr bool kStorm = F & kFeatStorm, kSources = F & kFeatSources;
    constexpr bool kAdaptive = F & kFeatAdaptive, kMap = F & kFeatMap;
    const int lane = threadIdx.x;

    // block -> (scenario, block within it).  The hardware deals consecutive
    // blocks to the 8 XCDs in turn; the remap gives each XCD one contiguous
    // run of the block list, so neighbouring tiles share an L2.
    int b = blockIdx.x;
    {
        const int nx = 8, per = nblocks / nx, rem = nblocks % nx;
        const int x = b % nx, i = b / nx;
        b = x * per + (x < rem ? x : rem) + i;
    }
    int sid = 0, tile = b;
    if (tiles) {
        const TileRef t = tiles[b];
        sid = t.sid;
        tile = t.local;
    }
    // the descriptor is read in place (no local copy: its uniform fields stay
    // reloadable scalar loads instead of pinned registers)
    const Scenario &d = scen[sid];
    const StepCommon &s = d.s;
    // iteration k of this scenario: k mod 6 from the launch phase, k itself
    // from the clock slot the previous iteration wrote
    const int k6 = (phase - d.phase0 + 6) % 6;
    Iter it;
    it.slot_a = k6 % 3; it.slot_b = (k6 + 1) % 3; it.slot_c = (k6 + 2) % 3;
    it.part_out = d.part + (size_t)(k6 & 1) * s.nparts;
    it.part_in = d.part + (size_t)((k6 + 1) & 1) * s.nparts;
    it.step_index = s.clk[it.slot_a].mstep >> 1;
    const StepTime st = step_time(s, it);
    if (d.status) {
        // batch member: a finished scenario leaves its blocks idle
        if (d.status->done) return;
        if (!(it.step_index < d.k_end && st.t + st.dt <= d.t_end)) {
            if (tile == d.ntiles && lane == 0) d.status->done = 1;
            return;
        }
    }
    const int ib = k6 & 1;
    // the state arrays (buffer resources in the FP32 form, see Arr); the
    // Manning class bytes, read only with a map, through a global pointer
    constexpr bool kBuf = sizeof(T) == 4 && sizeof(TH) == 4;
    const Arr<TH, kBuf> h_out(d.st[ib ^ 1][0]);
    const Arr<T, kBuf> um_out(d.st[ib ^ 1][1]), vn_out(d.st[ib ^ 1][2]);
    const gptr<const uint8_t> clsg = (gptr<const uint8_t>)d.cls;

    if (tile >= d.ntiles) {
        PitchedState<T, TH> ps{(const TH *)d.st[ib][0], (const T *)d.st[ib][1],
                               (const T *)d.st[ib][2], s.R, s.C, d.P, (TH)s.th};
        auto vel = [&](int r, int j, double &hh, double &ua, double &va) {
            hh = (double)ps.hv(r, j);
            ua = (double)ps.uua(r, j);
            va = (double)ps.vva(r, j);
        };
        const int part = tile - d.ntiles;
        bookkeeping(s, it, st, vel, part, lane);
        if (part == 0 && lane == 0 && d.status) {
            const long long kn = it.step_index + 1;
            const double tn = st.tm + st.dt;
            d.status->k = kn;
            d.status->t = tn;
            // the flag the kernel reads (done) is only set by a later launch;
            // fin tells the host now that nothing more will run
            if (kn >= d.k_end || (s.fixed_dt && !(tn + s.dt_fixed <= d.t_end))) d.status->fin = 1;
        }
        return;
    }

    const int strip = tile % d.nstrips, chunk = tile / d.nstrips;
    const int R = s.R, C = s.C, P = d.P;
    const int j0 = strip * d.W;
    const int j = j0 - 3 + lane;
    const bool colin = (j >= 0) && (j < C);
    const bool colout = (lane >= 3) && (lane < 3 + d.W) && (j < C);
    const int r0 = chunk * d.H;
    const int r1 = (r0 + d.H < R) ? r0 + d.H : R;

    // CFL candidates only for an adaptive time step; a wave-uniform flag, so
    // a scenario runs the same instructions alone and in a mixed batch
    const bool adapt = kAdaptive && !s.fixed_dt;
    const bool storm_on = kStorm && s.storm_on;
    const bool map_on = kMap && s.ncls_on;
    // FP32 arithmetic: storm factors and source terms in FP32.  FP32 depths:
    // water-surface tests from bed differences carried per face (z stays
    // FP64 only to form them) and stochastically rounded depth updates.
    constexpr bool kF32 = sizeof(T) == 4;
    constexpr bool kH32 = sizeof(TH) == 4;
    const T th = (T)s.th, gg = (T)s.g, dx = (T)s.dx, dy = (T)s.dy;
    const TH thh = (TH)s.th;
    const T rdx = (T)(1.0 / s.dx), rdy = (T)(1.0 / s.dy);
    const T dt2 = (T)st.dt2;
    const TH dt2h = (TH)st.dt2;
    const T dt2dx = (T)(st.dt2 / s.dx), dt2dy = (T)(st.dt2 / s.dy);
    // FP32 momentum: the upwind differences carry sums of two face
    // velocities, so dt2/dx and dt2/dy enter with the two halvings folded
    const T dt2dx4 = T(0.25) * dt2dx, dt2dy4 = T(0.25) * dt2dy, dt2half = T(0.5) * dt2;
    const T rn = (T)s.rn;

    // forcing of this iteration: first continuity at t, second at t + dt
    const double rr1 = rain_rate(s, st.t), rr2 = rain_rate(s, st.tm);
    const double ev1 = evap_rate(s, st.t), ev2 = evap_rate(s, st.tm);
    const bool storm1 = storm_live(s, st.t), storm2 = storm_live(s, st.tm);
    double gx1 = 0, gx2 = 0;
    T pk1 = 0, pk2 = 0;  // FP32 form: peak * gx of this column, 0 outside the storm window
    __shared__ T gy_s[2][kMaxTileRows + 8];
    if (storm_on) {
        const double x = ((double)j + 0.5) * s.dx;
        gx1 = storm_factor(x - (s.st_x0 + s.st_vx * st.t), s.st_sigma);
        gx2 = storm_factor(x - (s.st_x0 + s.st_vx * st.tm), s.st_sigma);
        pk1 = storm1 ? (T)(s.st_peak * gx1) : T(0);
        pk2 = storm2 ? (T)(s.st_peak * gx2) : T(0);
        for (int i = lane; i < d.H + 6; i += 64) {
            const double y1 = ((double)(r0 - 5 + i) + 0.5) * s.dy;
            const double y2 = ((double)(r0 - 6 + i) + 0.5) * s.dy;
            gy_s[0][i] = (T)storm_factor(y1 - (s.st_y0 + s.st_vy * st.t), s.st_sigma);
            gy_s[1][i] = (T)storm_factor(y2 - (s.st_y0 + s.st_vy * st.tm), s.st_sigma);
        }
        __syncthreads();
    }
    const T rr1t = (T)rr1, rr2t = (T)rr2, ev1t = (T)ev1, ev2t = (T)ev2;
    const T P_lake = (T)s.P_lake, P_land = (T)s.P_land;
    const TH ewl = (TH)s.EV_WL_min;

    // outlet
    const bool is_out_col = s.out_edge && (j == s.out_c);
    T qo_eff = 0;

    // carried rows: suffix 1 = row k-1, 2 = k-2, 3 = k-3
    TH h1 = 0, h2 = 0, h3 = 0;
    double z1 = 0, z2 = 0, z1w = 0;
    bool a1 = false, a2 = false, a3 = false, a1w = false;
    T u1 = 0, u2 = 0, v1 = 0, v2 = 0;
    T uu1_1 = 0, uu1_2 = 0;
    T vv1_1 = 0, vv1_2 = 0;
    T vva1 = 0;
    T dzn1 = 0, dzw1 = 0;  // FP32 form: z(k-2) - z(k-1) and z(k-1, j-1) - z(k-1, j)
    T nc1 = rn, nc2 = rn;
    T umnP = 0, vnnP = 0, umpP = 0, vnpP = 0;
    bool dryP = false;

    double acc_vol = 0, acc_src = 0, acc_clip = 0, acc_q = 0;
    // CFL: the largest |u| + sqrt(g h) per direction; min_f dx / s_f = dx / max_f s_f
    // because rounding dx / s is monotone in s
    T smax_x = 0, smax_y = 0;

    // software pipeline: row k+1 is loaded while row k is processed.  Rows
    // outside the grid load a clamped row and are then replaced by the ghost
    // values (0, inactive), so the loads need no branch.
    const int kbeg = r0 - 3, kend = r1 + 3;
    const int jc = colin ? j : (j < 0 ? 0 : C - 1);
    const Arr<const TH, kBuf> h_in(d.st[ib][0]);
    const Arr<const T, kBuf> um_in(d.st[ib][1]), vn_in(d.st[ib][2]);
    // the bed: FP64 elevations, or in the FP32 form the cell's bed-step pair
    // (8 bytes as well; read as one double's bits).  Ghosts read NaN either way.
    const Arr<const double, kBuf> zg(kH32 ? d.dz : d.z);
    const double zghost = kH32 ? __longlong_as_double(0x7FC000007FC00000LL) : __builtin_nan("");
    auto load_row = [&](int k, TH &hh, T &uu, T &vv, double &zz, uint8_t &cc) {
        const bool in = colin && k >= 0 && k < R;
        const int kc = k < 0 ? 0 : (k >= R ? R - 1 : k);
        const unsigned off = (unsigned)(kc * P + jc);
        const TH lh = h_in.load(off);
        const T lu = um_in.load(off), lv = vn_in.load(off);
        const double lz = zg.load(off);
        hh = in ? lh : TH(0);
        uu = in ? lu : T(0);
        vv = in ? lv : T(0);
        zz = in ? lz : zghost;
        cc = (map_on && in) ? clsg[off] : (uint8_t)0;
    };
    TH nh;
    T nu, nv;
    double nz;
    uint8_t ncl;
    load_row(kbeg, nh, nu, nv, nz, ncl);

    for (int k = kbeg; k < kend; ++k) {
        const TH h0 = nh;
        const T u0 = nu, v0 = nv;
        const double z0 = nz;
        const uint8_t cl0 = ncl;
        load_row(k + 1, nh, nu, nv, nz, ncl);
        // FP32 form: z0 holds the bed steps of row k's cell to its north (low
        // word) and west (high word) neighbours, NaN for an inactive cell
        const u32x2 zsteps = __builtin_bit_cast(u32x2, z0);
        const T zstep_n = (T)__uint_as_float(zsteps.x), zstep_w = (T)__uint_as_float(zsteps.y);
        // inactive cells and ghosts carry NaN
        const bool a0 = kH32 ? (zstep_w == zstep_w) : (z0 == z0);
        // west/east neighbours of rows k-1 and k-2, made here rather than
        // carried from the row before (shorter register live ranges)
        const TH h1w = from_west(h1);
        const T u1e = from_east(u1), u1w = from_west(u1);
        const T v1e = from_east(v1), v1w = from_west(v1);
        const T uu1_1w = from_west(uu1_1);
        const T vv1_1e = from_east(vv1_1), vv1_2e = from_east(vv1_2), vv1_1w = from_west(vv1_1);
        const T nc1w = map_on ? from_west(nc1) : rn;
        const T nc0 = map_on ? (T)s.ntab[cl0] : rn;

        // ---- row k: face depths and velocities (hm_hn, uu1_vv1) ----
        const TH h0w = from_west(h0);
        const double z0w = from_west(z0);
        const bool a0w = west_bit(a0, lane);
        T dzn0 = 0, dzw0 = 0;
        if constexpr (kH32) {
            dzn0 = zstep_n;  // bed step across the north face of row k: (float)(z(k-1) - z(k))
            dzw0 = zstep_w;  // bed step across the west face of row k: (float)(z(j-1) - z(j))
        }
        const TH hn0d = (j == 0) ? h0 : TH(0.50) * (h0 + h0w);
        const TH hm0d = (k == 0) ? h0 : ((k == R) ? h1 : TH(0.50) * (h0 + h1));
        const T hn0 = (T)hn0d, hm0 = (T)hm0d;
        const T uu1_0 = (hm0d >= thh) ? M::div_(u0, hm0) : T(0.00);
        const T vv1_0 = (hn0d >= thh) ? M::div_(v0, hn0) : T(0.0);
        const T vv1_0e = from_east(vv1_0), vv1_0w = from_west(vv1_0);
        const T uu1_0w = from_west(uu1_0);
        // uua_vva (489-522): vva of row k, uua of row k-1
        const T vva0 = (k > 0) ? T(0.250) * (vv1_0 + vv1_0e + vv1_1 + vv1_1e)
                               : T(0.50) * (vv1_0 + vv1_0e);
        const T uua1 = (j > 0) ? T(0.250) * (uu1_1 + uu1_0 + uu1_1w + uu1_0w)
                               : T(0.50) * (uu1_1 + uu1_0);

        if (adapt && colout && k >= r0 && k < r1) {
            // CFL candidates of the incoming state on the faces of row k
            if (k >= 1 && a0 && a1 && hm0d > thh) {
                const T w = M::abs_(uu1_0) + M::sqrt_(gg * hm0);
                smax_x = (w > smax_x) ? w : smax_x;
            }
            if (j >= 1 && a0 && a0w && hn0d > thh) {
                const T w = M::abs_(vv1_0) + M::sqrt_(gg * hn0);
                smax_y = (w > smax_y) ? w : smax_y;
            }
        }

        // ---- fluxes of row f = k-1 (flux, SW2D-EFDC.cu:97-353) ----
        const int f = k - 1;
        T umn = 0, vnn = 0;
        // rows before r0 - 1 feed no output row: the warm-up only builds the window
        const bool flux_um = k >= r0, flux_vn = k >= r0 && k <= r1 + 1;
        // x direction: um at the north face of (f, j); w = row f-1, e = row f.
        // Dry-bed branches B1-B3 all read (coef * hv) * sqrt(c2 * hv); each of
        // that and B4 (the momentum update) runs only when a lane needs it.
        const bool act_x = flux_um && f > 0 && f < R && a1 && a2 && ((h2 > thh) || (h1 > thh));
        if (__any(act_x)) {
            // face depth of row k-1, derived as hm0 was for it
            const TH hm1d = (k - 1 == 0) ? h1 : ((k - 1 == R) ? h2 : TH(0.50) * (h1 + h2));
            const T hm1 = (T)hm1d;
            const bool fr_x1 = !(hm1d <= thh);  // friction on: face depth above th
            // water surfaces: FP64 (the reference's arithmetic) or, in the
            // FP32 form, bed difference dz = z2 - z1 plus FP32 depths
            bool b1, b2;
            T dsurf, hv3;  // (h1 + z1) - (h2 + z2); max(surfaces) - max(beds)
            if constexpr (kH32) {
                const T dz = dzn1;
                b1 = h1 < dz;
                b2 = !b1 && (h2 < -dz);
                dsurf = (h1 - h2) - dz;
                hv3 = (dz <= T(0)) ? M::max_(h1, h2 + dz) : M::max_(h1 - dz, h2);
            } else {
                const double hhe = h1 + z1, hhw = h2 + z2;
                b1 = hhe < z2;
                b2 = !b1 && (hhw < z1);
                dsurf = (T)(hhe - h2 - z2);
                hv3 = (T)(__builtin_fmax(hhe, hhw) - __builtin_fmax(z1, z2));
            }
            const TH hhep = h1 - thh, hhwp = h2 - thh;
            const bool b3 = !b1 && !b2 && (hhep * hhwp < 0);
            const bool b4 = act_x && !b1 && !b2 && !b3;
            T coef = T(0.5440), hv = (T)h2, c2 = gg;
            bool wet = h2 > thh;
            if (b2) { coef = T(-0.544); hv = (T)h1; wet = h1 > thh; }
            if (b3) {
                const T sgnm = (T)MH::sign_(hhep - hhwp);  // hhan / |hhan|
                coef = -sgnm * T(0.350);
                hv = hv3;
                c2 = T(2.00) * gg;
                wet = (h1 > 0) || (h2 > 0);
            }
            T val = 0;
            if (__any(act_x && !b4)) val = wet ? coef * hv * M::sqrt_(c2 * hv) : T(0);
            if (__any(b4)) {
                const T u13 = gg * hm1 * dt2dx * dsurf;
                T u11, u12;
                if constexpr (kF32) {
                    // the halvings of the face velocities and of the upwind
                    // products folded into dt2/(4 dx) and dt2/(4 dy)
                    const T sr = uu1_0 + uu1_1, sl = uu1_1 + uu1_2;
                    const T pr = sr * (u0 + u1) + M::abs_(sr) * (u1 - u0);
                    const T pl = sl * (u1 + u2) + M::abs_(sl) * (u2 - u1);
                    u11 = dt2dx4 * (pr - pl);
                    const T su = vv1_1e + vv1_2e, sd = vv1_1 + vv1_2;
                    const T pu = su * (u1e + u1) + M::abs_(su) * (u1 - u1e);
                    const T pd = sd * (u1w + u1) + M::abs_(sd) * (u1w - u1);
                    u12 = (j != 0) ? dt2dy4 * (pu - pd) : T(0);
                } else {
                    const T ur = T(0.50) * (uu1_0 + uu1_1);
                    const T ul = T(0.50) * (uu1_1 + uu1_2);
                    const T umr = ur * (u0 + u1) * T(0.50) + M::abs_(ur) * (u1 - u0) * T(0.50);
                    const T uml = ul * (u1 + u2) * T(0.50) + M::abs_(ul) * (u2 - u1) * T(0.50);
                    u11 = dt2dx * (umr - uml);
                    const T vu = T(0.50) * (vv1_1e + vv1_2e);
                    const T vd = T(0.50) * (vv1_1 + vv1_2);
                    const T umu = vu * (u1e + u1) * T(0.50) + M::abs_(vu) * (u1 - u1e) * T(0.50);
                    const T umd = vd * (u1w + u1) * T(0.50) + M::abs_(vd) * (u1w - u1) * T(0.50);
                    u12 = (j != 0) ? dt2dy * (umu - umd) : T(0);
                }
                const T nf = map_on ? T(0.5) * (nc2 + nc1) : rn;
                T ram = 0.00;
                if (b4 && fr_x1) {
                    const T sqx = M::sqrt_(uu1_1 * uu1_1 + vva1 * vva1);
                    ram = M::div_(gg * nf * nf * sqx, pow1333(hm1));
                }
                T up;
                if constexpr (kF32) {
                    const T a = dt2half * ram;
                    up = M::div_((T(1.00) - a) * u1 + (-u11 - u12 - u13), T(1.00) + a);
                } else {
                    up = M::div_((T(1.00) - dt2 * ram * T(0.50)) * u1 + (-u11 - u12 - u13),
                                 T(1.00) + dt2 * ram * T(0.50));
                }
                if (b4) val = up;
            }
            umn = act_x ? val : T(0);
        }
        // y direction: vn at the west face of (f, j); s = column j-1, n = column j
        const bool act_y = flux_vn && j > 0 && j < C && a1 && a1w && ((h1w > thh) || (h1 > thh));
        if (__any(act_y)) {
            const TH hn1d = (j == 0) ? h1 : TH(0.50) * (h1 + h1w);
            const T hn1 = (T)hn1d;
            const bool fr_y1 = !(hn1d <= thh);
            bool b1, b2;
            T dsurf, hv3;  // (h1 + z1) - (h1w + z1w); max(surfaces) - max(beds)
            if constexpr (kH32) {
                const T dz = dzw1;
                b1 = h1 < dz;
                b2 = !b1 && (h1w < -dz);
                dsurf = (h1 - h1w) - dz;
                hv3 = (dz <= T(0)) ? M::max_(h1, h1w + dz) : M::max_(h1 - dz, h1w);
            } else {
                const double hhn = h1 + z1, hhs = h1w + z1w;
                b1 = hhn < z1w;
                b2 = !b1 && (hhs < z1);
                dsurf = (T)(hhn - h1w - z1w);
                hv3 = (T)(__builtin_fmax(hhn, hhs) - __builtin_fmax(z1, z1w));
            }
            const TH hhnp = h1 - thh, hhsp = h1w - thh;
            const bool b3 = !b1 && !b2 && (hhnp * hhsp < 0);
            const bool b4 = act_y && !b1 && !b2 && !b3;
            T coef = T(0.5440), hv = (T)h1w, c2 = gg;
            bool wet = h1w > thh;
            if (b2) { coef = T(-0.544); hv = (T)h1; wet = h1 > thh; }
            if (b3) {
                const T sgnm = (T)MH::sign_(hhnp - hhsp);
                coef = -sgnm * T(0.350);
                hv = hv3;
                c2 = T(2.00) * gg;
                wet = (h1 > 0) || (h1w > 0);
            }
            T val = 0;
            if (__any(act_y && !b4)) val = wet ? coef * hv * M::sqrt_(c2 * hv) : T(0);
            if (__any(b4)) {
                const T v13 = gg * hn1 * dt2dy * dsurf;
                T v11, v12;
                if constexpr (kF32) {
                    const T sa = uu1_0 + uu1_0w, sb = uu1_1 + uu1_1w;
                    const T pr = sa * (v0 + v1) + M::abs_(sa) * (v1 - v0);
                    const T pl = sb * (v1 + v2) + M::abs_(sb) * (v2 - v1);
                    v11 = (f != 0 && f != R) ? dt2dx4 * (pr - pl) : T(0);
                    const T sp = vv1_1e + vv1_1, sq = vv1_1 + vv1_1w;
                    const T pu = sp * (v1e + v1) + M::abs_(sp) * (v1 - v1e);
                    const T pd = sq * (v1w + v1) + M::abs_(sq) * (v1w - v1);
                    v12 = dt2dy4 * (pu - pd);
                } else {
                    const T av = T(0.50) * (uu1_0 + uu1_0w);
                    const T bv = T(0.50) * (uu1_1 + uu1_1w);
                    const T vnr = av * (v0 + v1) * T(0.50) + M::abs_(av) * (v1 - v0) * T(0.50);
                    const T vnl = bv * (v1 + v2) * T(0.50) + M::abs_(bv) * (v2 - v1) * T(0.50);
                    v11 = (f != 0 && f != R) ? dt2dx * (vnr - vnl) : T(0);
                    const T pv = T(0.50) * (vv1_1e + vv1_1);
                    const T qv = T(0.50) * (vv1_1 + vv1_1w);
                    const T vnu = pv * (v1e + v1) * T(0.50) + M::abs_(pv) * (v1 - v1e) * T(0.50);
                    const T vnd = qv * (v1w + v1) * T(0.50) + M::abs_(qv) * (v1w - v1) * T(0.50);
                    v12 = dt2dy * (vnu - vnd);
                }
                const T nf = map_on ? T(0.5) * (nc1w + nc1) : rn;
                T ram = 0.00;
                if (b4 && fr_y1) {
                    const T sqy = M::sqrt_(uua1 * uua1 + vv1_1 * vv1_1);
                    ram = M::div_(gg * nf * nf * sqy, pow1333(hn1));
                }
                T up;
                if constexpr (kF32) {
                    const T a = dt2half * ram;
                    up = M::div_((T(1.00) - a) * v1 + (-v11 - v12 - v13), T(1.00) + a);
                } else {
                    up = M::div_((T(1.00) - dt2 * ram * T(0.50)) * v1 + (-v11 - v12 - v13),
                                 T(1.00) + dt2 * ram * T(0.50));
                }
                if (b4) val = up;
            }
            vnn = act_y ? val : T(0);
        }

        // ---- first continuity of row g = k-2 at time t (602-664) ----
        const int g = k - 2;
        bool dryg = false;
        T umpg = 0, vnpg = 0;
        if (k >= r0 + 1) {
            const T vnnPe = from_east(vnnP);
            const bool wet = h2 > ewl;
            const T Pc = wet ? P_lake : P_land;
            const T E = wet ? ev1t : T(0.000);
            T rr = rr1t;
            if (storm_on) {
                const int gi = g - (r0 - 5);
                const T gy = gy_s[0][gi < 0 ? 0 : gi];
                if constexpr (kF32) rr = pk1 * gy;
                else rr = storm1 ? (T)(s.st_peak * (gx1 * gy)) : T(0);
            }
            T ql = 0;
            for (int i = 0; kSources && i < s.nst; ++i)
                if (s.src_r[i] == g && s.src_c[i] == j) ql = (T)source_rate(s, i, st.t);
            T ums = umn, umnn = umnP, vne = vnnPe, vnw = vnnP;
            if (is_out_col && g == s.out_r) {
                const T h2t = (T)h2;
                const T qo = (h2 > thh) ? T(0.5440) * h2t * M::sqrt_(gg * h2t) : T(0);
                if (s.out_edge == 1) ums = qo;
                else if (s.out_edge == 2) umnn = -qo;
                else if (s.out_edge == 3) vnw = -qo;
                else vne = qo;
                qo_eff = qo;
            }
            const T rhs = M::divc_(ums - umnn, dx, rdx) + M::divc_(vne - vnw, dy, rdy) - ql -
                          rr * Pc + E;
            const TH hr = h2 - dt2h * (TH)rhs;
            TH h1v = MH::max_(hr, TH(0.00));
            if (!a2) h1v = 0.00;
            dryg = a2 && (h1v < thh);
            // ---- treat_error on the faces of row g (695-725) ----
            umpg = umnP;
            if (dryP && umpg > 0) umpg = 0;
            if (dryg && umpg < 0) umpg = 0;
            const bool drygw = west_bit(dryg, lane);
            vnpg = vnnP;
            if (drygw && vnpg > 0) vnpg = 0;
            if (dryg && vnpg < 0) vnpg = 0;
            if (is_out_col && g == s.out_r && dryg) qo_eff = 0;
        }

        // ---- second continuity of row q = k-3 at time t + dt; store ----
        const int q = k - 3;
        if (k >= r0 + 3) {
            const T vnpPe = from_east(vnpP);
            const bool wet = h3 > ewl;
            const T Pc = wet ? P_lake : P_land;
            const T E = wet ? ev2t : T(0.000);
            T rr = rr2t;
            if (storm_on) {
                const int qi = q - (r0 - 6);
                const T gy = gy_s[1][qi < 0 ? 0 : qi];
                if constexpr (kF32) rr = pk2 * gy;
                else rr = storm2 ? (T)(s.st_peak * (gx2 * gy)) : T(0);
            }
            T ql = 0;
            for (int i = 0; kSources && i < s.nst; ++i)
                if (s.src_r[i] == q && s.src_c[i] == j) ql = (T)source_rate(s, i, st.tm);
            T ums = umpg, umnn = umpP, vne = vnpPe, vnw = vnpP;
            const bool outcell = is_out_col && q == s.out_r;
            if (outcell) {
                if (s.out_edge == 1) ums = qo_eff;
                else if (s.out_edge == 2) umnn = -qo_eff;
                else if (s.out_edge == 3) vnw = -qo_eff;
                else vne = qo_eff;
            }
            const T rhs = M::divc_(ums - umnn, dx, rdx) + M::divc_(vne - vnw, dy, rdy) - ql -
                          rr * Pc + E;
            TH h2v;
            double clip = 0.0;
            if constexpr (kH32) {
                // FP32 form: the depth update h3 - dt2 rhs is rounded
                // stochastically.  TwoSum gives the exact rounding error er of
                // sm = fl(h3 + tt); the result moves to the FP32 neighbour of sm
                // on er's side with probability |er| / spacing, so rounding
                // errors average out instead of adding up coherently (every
                // cell of a uniformly wetted slope rounds the same small rain
                // increment the same way).  The random number is a fixed hash
                // of (cell, iteration): results are reproducible and do not
                // depend on the tiling or on other scenarios.
                const T tt = -dt2 * rhs;
                const T sm = h3 + tt;
                const T bp = sm - h3;
                const T er = (h3 - (sm - bp)) + (tt - bp);
                h2v = sm;
                if (sm > T(0) && er != T(0)) {
                    const uint32_t bits = __float_as_uint(sm);
                    const T nb = __uint_as_float(er > T(0) ? bits + 1u : bits - 1u);
                    const T pr = er * __builtin_amdgcn_rcpf(nb - sm);
                    if (cell_uniform((uint32_t)(q * C + j), (uint32_t)it.step_index) < pr) h2v = nb;
                }
                // the exact update sm + er is negative iff sm < 0, or sm = 0 and
                // er < 0 (|er| is at most half an ulp of sm); its FP64 value
                // (exact: 48 bits suffice) is needed only then
                if (sm < T(0) || (sm == T(0) && er < T(0))) {
                    clip = -((double)sm + (double)er);
                    h2v = 0;
                }
            } else {
                const TH hr = h3 - dt2h * (TH)rhs;
                h2v = MH::max_(hr, TH(0.00));
                if (hr < 0) clip = -(double)hr;
            }
            if (!a3) h2v = 0.00;
            if (colout && q >= r0 && q < r1) {
                const unsigned off = (unsigned)(q * P + j);
                h_out.store(off, h2v);
                um_out.store(off, umpP);
                vn_out.store(off, vnpP);
                if (a3) {
                    acc_vol += (double)h2v;
                    if constexpr (kF32) acc_src += (double)(ql + rr * Pc - E);  // x dt2 below
                    else acc_src += st.dt2 * ((double)ql + (double)rr * (double)Pc - (double)E);
                    acc_clip += clip;
                }
                if (outcell) acc_q += (double)qo_eff * s.dy;
            }
        }

        // ---- rotate the row window ----
        h3 = h2; h2 = h1; h1 = h0;
        z2 = z1; z1 = z0; z1w = z0w;
        a3 = a2; a2 = a1; a1 = a0; a1w = a0w;
        u2 = u1; u1 = u0;
        v2 = v1; v1 = v0;
        uu1_2 = uu1_1; uu1_1 = uu1_0;
        vv1_2 = vv1_1; vv1_1 = vv1_0;
        vva1 = vva0;
        dzn1 = dzn0; dzw1 = dzw0;
        nc2 = nc1; nc1 = nc0;
        umnP = umn; vnnP = vnn;
        dryP = dryg; umpP = umpg; vnpP = vnpg;
    }

    // ---- per-tile partial sums (fixed-order wave reductions) ----
    if constexpr (kF32) acc_src *= st.dt2;
    acc_vol = wave_sum(acc_vol);
    acc_src = wave_sum(acc_src);
    acc_clip = wave_sum(acc_clip);
    acc_q = wave_sum(acc_q);
    double cfl = INFINITY;
    if (adapt) {
        const double mx = wave_max((double)smax_x), my = wave_max((double)smax_y);
        const double wx = mx > 0.0 ? s.dx / mx : INFINITY;
        const double wy = my > 0.0 ? s.dy / my : INFINITY;
        cfl = (wx < wy) ? wx : wy;
    }
    if (lane == 0) {
        Partial p;
        p.vol = acc_vol; p.src = acc_src; p.clip = acc_clip; p.qout = acc_q;
        it.part_out[tile] = p;
        if (adapt && cfl < INFINITY)
            atomicMin(&s.clk[it.slot_b].cfl_bits, (unsigned long long)__double_as_longlong(cfl));
    }
}

// initial CFL minimum of the state (adaptive mode)
template <typename T, typename TH>
__global__ void __launch_bounds__(256) init_cfl(const TH *h, const T *um, const T *vn,
                                                const double *z, int R, int C, int P, double dx,
                                                double dy, double g, double th, Clock *clk)
{
    using M = num<T>;
    const long n = (long)R * C;
    double cfl = INFINITY;
    for (long idx = (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x) {
        const int r = (int)(idx / C), j = (int)(idx % C);
        const long o = (long)r * P + j;
        const bool ac = (z[o] == z[o]);
        if (r >= 1) {
            const bool an = (z[o - P] == z[o - P]);
            const TH hm = TH(0.50) * (h[o] + h[o - P]);
            if (ac && an && hm > (TH)th) {
                const T uu1 = (hm >= (TH)th) ? M::div_(um[o], (T)hm) : T(0.00);
                const double w = dx / ((double)M::abs_(uu1) + sqrt(g * (double)hm));
                cfl = (w < cfl) ? w : cfl;
            }
        }
        if (j >= 1) {
            const bool aw = (z[o - 1] == z[o - 1]);
            const TH hn = TH(0.50) * (h[o] + h[o - 1]);
            if (ac && aw && hn > (TH)th) {
                const T vv1 = (hn >= (TH)th) ? M::div_(vn[o], (T)hn) : T(0.0);
                const double w = dy / ((double)M::abs_(vv1) + sqrt(g * (double)hn));
                cfl = (w < cfl) ? w : cfl;
            }
        }
    }
    cfl = wave_min(cfl);
    if ((threadIdx.x & 63) == 0 && cfl < INFINITY)
        atomicMin(&clk[0].cfl_bits, (unsigned long long)__double_as_longlong(cfl));
}

// one derived field of the current state (hm_hn, uu1_vv1, uu_vv, uua_vva) over
// the cells first .. first + count - 1 of the row-major rows x cols order;
// out[i] receives cell first + i
template <typename T, typename TH>
__global__ void __launch_bounds__(256) derive_field(const TH *h, const T *um, const T *vn,
                                                    const double *z, int R, int C, int P,
                                                    double th, int field, double *out,
                                                    long first, long count)
{
    PitchedState<T, TH> ps{h, um, vn, R, C, P, (TH)th};
    const long n = first + count;
    for (long idx = first + (long)blockIdx.x * blockDim.x + threadIdx.x; idx < n;
         idx += (long)gridDim.x * blockDim.x) {
        const int r = (int)(idx / C), j = (int)(idx % C);
        const long o = (long)r * P + j;
        double v = 0;
        switch (field) {
        case SW2D_FIELD_H: v = h[o]; break;
        case SW2D_FIELD_UM: v = um[o]; break;
        case SW2D_FIELD_VN: v = vn[o]; break;
        case SW2D_FIELD_HM: v = (r == 0) ? h[o] : TH(0.50) * (h[o] + h[o - P]); break;
        case SW2D_FIELD_HN: v = (j == 0) ? h[o] : TH(0.50) * (h[o] + h[o - 1]); break;
        case SW2D_FIELD_UU1: v = ps.uu1(r, j); break;
        case SW2D_FIELD_VV1: v = ps.vv1(r, j); break;
        case SW2D_FIELD_UU: v = (h[o] >= (TH)th) ? (ps.uu1(r + 1, j) + ps.uu1(r, j)) / T(2.00) : T(0.00); break;
        case SW2D_FIELD_VV: v = (h[o] >= (TH)th) ? (ps.vv1(r, j + 1) + ps.vv1(r, j)) / T(2.00) : T(0.00); break;
        case SW2D_FIELD_UUA: v = ps.uua(r, j); break;
        case SW2D_FIELD_VVA: v = ps.vva(r, j); break;
        case SW2D_FIELD_Z: v = z[o]; break;
        default: break;
        }
        out[idx - first] = v;
    }
}

// lds: dynamic LDS bytes reserved per block and left unused (0: none); a
// batch uses it to cap the resident waves per SIMD
template <typename T, typename TH>
static hipError_t launch_fused_impl(hipStream_t stream, const Scenario *scen, const TileRef *tiles,
                                    int nblocks, int phase, int features, size_t lds)
{
    switch (fused_variant(features)) {
    case kVariant0:
        hipLaunchKernelGGL((fused_step<T, TH, kVariant0>), dim3(nblocks), dim3(64), lds, stream, scen, tiles, nblocks, phase);
        break;
    case kVariant1:
        hipLaunchKernelGGL((fused_step<T, TH, kVariant1>), dim3(nblocks), dim3(64), lds, stream, scen, tiles, nblocks, phase);
        break;
    case kVariant2:
        hipLaunchKernelGGL((fused_step<T, TH, kVariant2>), dim3(nblocks), dim3(64), lds, stream, scen, tiles, nblocks, phase);
        break;
    case kVariant3:
        hipLaunchKernelGGL((fused_step<T, TH, kVariant3>), dim3(nblocks), dim3(64), lds, stream, scen, tiles, nblocks, phase);
        break;
    default:
        hipLaunchKernelGGL((fused_step<T, TH, kFeatAll>), dim3(nblocks), dim3(64), lds, stream, scen, tiles, nblocks, phase);
        break;
    }
    return hipGetLastError();
}

// resident one-wave blocks per CU of the instantiation that runs `features`
template <typename T, typename TH>
static int fused_waves_per_cu_impl(int features)
{
    const void *fn;
    switch (fused_variant(features)) {
    case kVariant0: fn = (const void *)fused_step<T, TH, kVariant0>; break;
    case kVariant1: fn = (const void *)fused_step<T, TH, kVariant1>; break;
    case kVariant2: fn = (const void *)fused_step<T, TH, kVariant2>; break;
    case kVariant3: fn = (const void *)fused_step<T, TH, kVariant3>; break;
    default: fn = (const void *)fused_step<T, TH, kFeatAll>; break;
    }
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, 64, 0) != hipSuccess) return 0;
    return nb;
}

template <typename T, typename TH>
static hipError_t launch_init_impl(sw2d_handle *h)
{
    if (!h->prm.adaptive) return hipSuccess;
    const long n = (long)h->R * h->C;
    int blocks = (int)((n + 255) / 256);
    if (blocks > 4096) blocks = 4096;
    hipLaunchKernelGGL((init_cfl<T, TH>), dim3(blocks), dim3(256), 0, h->stream,
                       (const TH *)h->h[h->cur], (const T *)h->um[h->cur], (const T *)h->vn[h->cur],
                       (const double *)h->z, h->R, h->C, h->P, h->grid.cellsize, h->grid.cellsize,
                       h->prm.g, h->prm.th, h->clk);
    return hipGetLastError();
}

template <typename T, typename TH>
static hipError_t launch_derive_impl(sw2d_handle *h, int field, double *out_dev, long first,
                                     long count)
{
    int blocks = (int)((count + 255) / 256);
    if (blocks > 8192) blocks = 8192;
    if (blocks < 1) return hipSuccess;
    hipLaunchKernelGGL((derive_field<T, TH>), dim3(blocks), dim3(256), 0, h->stream,
                       (const TH *)h->h[h->cur], (const T *)h->um[h->cur], (const T *)h->vn[h->cur],
                       (const double *)h->z, h->R, h->C, h->P, h->prm.th, field, out_dev, first,
                       count);
    return hipGetLastError();
}

}  // namespace sw2d
