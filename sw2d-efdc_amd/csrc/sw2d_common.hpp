// sw2d_common.hpp -- device code shared by the fused and the plain kernels:
// iteration timing, forcing evaluation, wave reductions, the bookkeeping
// (clock ring, volume ledger, hydrograph, lake-inflow integration).
#pragma once
#include "sw2d_internal.hpp"
#include "sw2d_math.hpp"

namespace sw2d {

struct StepTime {
    double t, dt, dt2, tm;
};

// start time and length of this iteration from the clock ring
__device__ __forceinline__ StepTime step_time(const StepCommon &s, const Iter &it)
{
    const Clock &ck = s.clk[it.slot_a];
    StepTime st;
    st.t = ck.t;
    if (s.fixed_dt) {
        st.dt = s.dt_fixed;
        st.dt2 = s.dt2_fixed;
    } else {
        double m = __longlong_as_double((long long)ck.cfl_bits);
        double d = s.courant * m;
        st.dt2 = (d < s.dt2_max) ? d : s.dt2_max;
        st.dt = 0.5 * st.dt2;
    }
    st.tm = st.t + st.dt;
    return st;
}

// rr of stream_flow in m/s (SW2D-EFDC.cu:546-553); 0 without a rain series
__device__ __forceinline__ double rain_rate(const StepCommon &s, double t)
{
    if (s.nrain <= 0) return 0.0;
    double rr = series_value(s.rain, s.nrain, s.dtrain, t);
    return rr / (s.dtrain * 1000.0);
}

// Ev of evaporation_load in m/s (SW2D-EFDC.cu:588-600)
__device__ __forceinline__ double evap_rate(const StepCommon &s, double t)
{
    if (!s.evap_on) return 0.0;
    double e = series_value(s.evapo, s.nevapo, s.dtrain, t);
    return e / (s.dtrain * 1000.0);
}

// ql of source i in m/s (SW2D-EFDC.cu:531-539)
__device__ __forceinline__ double source_rate(const StepCommon &s, int i, double t)
{
    double ql = source_value(s.Q, s.nq, s.nst, i, s.dtoq, t);
    return ql / (s.dx * s.dy);
}

// uniform number in [0, 1) from (cell, iteration): a 32-bit integer hash
// (multiply-xorshift finaliser), 24 bits of it scaled by 2^-24
__device__ __forceinline__ float cell_uniform(uint32_t cell, uint32_t iter)
{
    uint32_t x = cell * 0x9E3779B1u ^ iter * 0x85EBCA77u;
    x ^= x >> 16;
    x *= 0x7FEB352Du;
    x ^= x >> 15;
    x *= 0x846CA68Bu;
    x ^= x >> 16;
    return (float)(x >> 8) * 5.9604644775390625e-08f;
}

// one factor of the moving Gaussian storm
__device__ __forceinline__ double storm_factor(double d, double sigma)
{
    return exp_d(-(d * d) / (2.0 * sigma * sigma));
}

__device__ __forceinline__ bool storm_live(const StepCommon &s, double t)
{
    return s.storm_on && (t >= s.st_t0) && (t < s.st_t1);
}

// wave-wide reductions with a fixed combination order (lane 0 holds the result)
__device__ __forceinline__ double wave_sum(double v)
{
    for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
    return v;
}
__device__ __forceinline__ double wave_max(double v)
{
    for (int off = 32; off >= 1; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = (o > v) ? o : v;
    }
    return v;
}
__device__ __forceinline__ double wave_min(double v)
{
    for (int off = 32; off >= 1; off >>= 1) {
        double o = __shfl_xor(v, off, 64);
        v = (o < v) ? o : v;
    }
    return v;
}

// inflow rate of one coupling record (couple_BC_efdc, SW2D-EFDC.cu:764-821)
__device__ __forceinline__ double bc_rate(int kind, double h, double uua, double vva, double dx,
                                          double dy)
{
    switch (kind) {
    case 0: return (uua < 0) ? -uua * h * dx : 0.00;
    case 1: return (uua > 0) ? uua * h * dx : 0.00;
    case 2: return (vva < 0) ? -vva * h * dy : 0.00;
    default: return (vva > 0) ? vva * h * dy : 0.00;
    }
}

// Face velocities of a state held in a pitched (rows x pitch) layout, with the
// reference's edge rules (hm_hn 362-398, uu1_vv1 404-452): every face outside
// the grid, on row 0 and on column 0 carries zero flux.  T is the type of the
// fluxes, TH that of the depths (the mixed form keeps depths in FP64): the
// face depth and its dry test are taken in TH, the velocity in T.
template <typename T, typename TH = T>
struct PitchedState {
    const TH *h;
    const T *um, *vn;
    int R, C, P;
    TH th;
    __device__ __forceinline__ TH hv(int r, int j) const { return h[(long)r * P + j]; }
    __device__ __forceinline__ T uu1(int r, int j) const
    {
        if (r <= 0 || r >= R || j < 0 || j >= C) return T(0);
        TH hm = TH(0.50) * (hv(r, j) + hv(r - 1, j));
        return (hm >= th) ? um[(long)r * P + j] / (T)hm : T(0.00);
    }
    __device__ __forceinline__ T vv1(int r, int j) const
    {
        if (r < 0 || r >= R || j <= 0 || j >= C) return T(0);
        TH hn = TH(0.50) * (hv(r, j) + hv(r, j - 1));
        return (hn >= th) ? vn[(long)r * P + j] / (T)hn : T(0.0);
    }
    // uua_vva, SW2D-EFDC.cu:489-522
    __device__ __forceinline__ T uua(int r, int j) const
    {
        if (j > 0) return T(0.250) * (uu1(r, j) + uu1(r + 1, j) + uu1(r, j - 1) + uu1(r + 1, j - 1));
        return T(0.50) * (uu1(r, j) + uu1(r + 1, j));
    }
    __device__ __forceinline__ T vva(int r, int j) const
    {
        if (r > 0) return T(0.250) * (vv1(r, j) + vv1(r, j + 1) + vv1(r - 1, j) + vv1(r - 1, j + 1));
        return T(0.50) * (vv1(r, j) + vv1(r, j + 1));
    }
};

// Bookkeeping of iteration k.  Part 0 (one wave): advance the clock ring,
// reduce the partial sums of iteration k-1 into the ledger (fixed order) and
// record its hydrograph sample.  Part b >= 1 (one wave each): integrate the
// lake inflow of records 64(b-1) .. 64b-1 over the state at the start of
// iteration k.  uua/vva are evaluated by the functor F.
__device__ __forceinline__ void reduce_partials(const Partial *p, int n, int lane, double &vol,
                                                double &src, double &clip, double &qo)
{
    vol = 0; src = 0; clip = 0; qo = 0;
    // four independent loads in flight per lane
    int i = lane;
    for (; i + 192 < n; i += 256) {
        const Partial a = p[i], b = p[i + 64], c = p[i + 128], d = p[i + 192];
        vol += a.vol; src += a.src; clip += a.clip; qo += a.qout;
        vol += b.vol; src += b.src; clip += b.clip; qo += b.qout;
        vol += c.vol; src += c.src; clip += c.clip; qo += c.qout;
        vol += d.vol; src += d.src; clip += d.clip; qo += d.qout;
    }
    for (; i < n; i += 64) {
        const Partial a = p[i];
        vol += a.vol; src += a.src; clip += a.clip; qo += a.qout;
    }
    vol = wave_sum(vol); src = wave_sum(src); clip = wave_sum(clip); qo = wave_sum(qo);
}

template <typename F>
__device__ void bookkeeping(const StepCommon &s, const Iter &it, const StepTime &st, const F &vel,
                            const int part,
                            const int lane)
{
    if (part > 0) {
        const int i = (part - 1) * 64 + lane;
        if (i < s.nbc) {
            const int r = s.bc_r[i], j = s.bc_c[i];
            double hh, ua, va;
            vel(r, j, hh, ua, va);
            s.bc_acc[i] += bc_rate(s.bc_kind[i], hh, ua, va, s.dx, s.dy) * st.dt2;
        }
        return;
    }
    if (lane == 0) {
        Clock &nb = s.clk[it.slot_b];
        nb.t = st.tm + st.dt;
        nb.dt2_prev = st.dt2;
        nb.mstep = s.clk[it.slot_a].mstep + 2;
        s.clk[it.slot_c].cfl_bits = kInfBits;
    }
    if (it.step_index > 0) {
        double vol, src, clip, qo;
        reduce_partials(it.part_in, s.nparts, lane, vol, src, clip, qo);
        if (lane == 0) {
            Ledger &L = *s.ledger;
            double dt2p = s.clk[it.slot_a].dt2_prev;
            L.v_src += src * s.dx * s.dy;
            L.v_clip += clip * s.dx * s.dy;
            L.v_out += qo * dt2p;
            L.q_last = qo;
            if (qo > L.q_peak) { L.q_peak = qo; L.t_peak = st.t; }
            L.vol_last = vol * s.dx * s.dy;
            L.reduced += 1;
            long long i = it.step_index - 1;
            if (s.hydro && i < s.hydro_cap) {
                s.hydro[2 * i] = st.t;
                s.hydro[2 * i + 1] = qo;
            }
        }
    }
}

// Ledger view for the host: the committed ledger plus the partials of the
// last iteration, which the next iteration would reduce.
__device__ __forceinline__ void ledger_peek(const StepCommon &s, Ledger *out, const int lane)
{
    const Iter it = iter_of(s);
    double vol = 0, src = 0, clip = 0, qo = 0;
    if (it.step_index > 0) reduce_partials(it.part_in, s.nparts, lane, vol, src, clip, qo);
    if (lane == 0) {
        Ledger L = *s.ledger;
        if (it.step_index > 0) {
            double dt2p = s.clk[it.slot_a].dt2_prev;
            double t = s.clk[it.slot_a].t;
            L.v_src += src * s.dx * s.dy;
            L.v_clip += clip * s.dx * s.dy;
            L.v_out += qo * dt2p;
            L.q_last = qo;
            if (qo > L.q_peak) { L.q_peak = qo; L.t_peak = t; }
            L.vol_last = vol * s.dx * s.dy;
            L.reduced += 1;
            long long i = it.step_index - 1;
            if (s.hydro && i < s.hydro_cap) {
                s.hydro[2 * i] = t;
                s.hydro[2 * i + 1] = qo;
            }
        }
        *out = L;
    }
}

}  // namespace sw2d
