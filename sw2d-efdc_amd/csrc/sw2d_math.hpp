// sw2d_math.hpp -- device math of the SW2D kernels.
//
// FP64: pow(x, 1.3333330) (the Manning depth exponent of SW2D-EFDC.cu:194,
// 311) and exp() are built from correctly rounded operations only (+ - * /,
// explicit fma, floor, exponent bit manipulation), so that the result is a
// fixed function of the input on every machine: the CPU FP64 check of this
// library evaluates the same sequence and must agree bit for bit.
// Translation units that include this header are compiled with
// -ffp-contract=off (no fusing beyond the fma written here).
//
// pow for 2^-700 <= x <= 2^500 (every depth the kernels raise to the
// power): table-driven log and exp, no division.  x = 2^e m, m in [1, 2);
// i = the top 6 bits of m's fraction, r = m RECIP[i] - 1 (one fma, |r| <
// 2^-7), ln x = e ln2 + NLOG[i] + log1p(r) with log1p by its Taylor series to
// r^7; t = 1.3333330 ln x; k = rint(64 t / ln2), r' = t - k ln2/64
// (two-constant ln2), exp(t) = 2^(k>>6) EXP2[k&63] (1 + P(r')) with P the
// Taylor series of exp(r') - 1 to r'^6.  The tables (sw2d_pow_tables.h) hold
// correctly rounded values.  Error against a correctly rounded pow: at most
// 20 ulp for x in [1e-5, 1e3] (measured on 2e7 points), dominated by the
// rounding of t.
//
// Other x, and exp()/log() themselves: x = 2^e m, m in (sqrt(1/2), sqrt(2)];
// log m = 2 atanh(s), s = (m-1)/(m+1), arranged as f - (hf - s (hf + R)),
// hf = f^2/2, R(z) = sum_k 2 z^k / (2k+1), z = s^2 (the atanh-series
// arrangement of Sun's fdlibm e_log.c with plain Taylor coefficients);
// exp(t) = 2^k exp(r), r = t - k ln2 with a two-constant (Cody-Waite) ln2,
// exp(r) by Taylor to r^13; both polynomials by Horner's rule with fma.
//
// FP32: the hardware log2/exp2 approximations (about 1 ulp each).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "sw2d_pow_tables.h"

namespace sw2d {

constexpr double kLn2Hi = 6.93147180369123816490e-01;
constexpr double kLn2Lo = 1.90821492927058770002e-10;
constexpr double kInvLn2 = 1.44269504088896338700e+00;
constexpr double kSqrt2 = 1.41421356237309514547e+00;
constexpr double kManningExp = 1.3333330;

__device__ __forceinline__ double log_pos(double x)
{
    uint64_t b = (uint64_t)__double_as_longlong(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double m = __longlong_as_double((long long)((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
    if (m > kSqrt2) { m = m * 0.5; e = e + 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    double R = 2.0 / 25.0;
    R = __builtin_fma(z, R, 2.0 / 23.0);
    R = __builtin_fma(z, R, 2.0 / 21.0);
    R = __builtin_fma(z, R, 2.0 / 19.0);
    R = __builtin_fma(z, R, 2.0 / 17.0);
    R = __builtin_fma(z, R, 2.0 / 15.0);
    R = __builtin_fma(z, R, 2.0 / 13.0);
    R = __builtin_fma(z, R, 2.0 / 11.0);
    R = __builtin_fma(z, R, 2.0 / 9.0);
    R = __builtin_fma(z, R, 2.0 / 7.0);
    R = __builtin_fma(z, R, 2.0 / 5.0);
    R = __builtin_fma(z, R, 2.0 / 3.0);
    R = z * R;
    double hf = 0.5 * f * f;
    double lm = f - (hf - s * (hf + R));
    double de = (double)e;
    return de * kLn2Hi + (de * kLn2Lo + lm);
}

__device__ __forceinline__ double exp_d(double t)
{
    if (t != t) return t;
    if (t > 709.0) return __longlong_as_double(0x7ff0000000000000LL);
    if (t < -708.0) return 0.0;
    double kd = floor(t * kInvLn2 + 0.5);
    int k = (int)kd;
    double r = (t - kd * kLn2Hi) - kd * kLn2Lo;
    double p = 1.0 / 6227020800.0;
    p = __builtin_fma(r, p, 1.0 / 479001600.0);
    p = __builtin_fma(r, p, 1.0 / 39916800.0);
    p = __builtin_fma(r, p, 1.0 / 3628800.0);
    p = __builtin_fma(r, p, 1.0 / 362880.0);
    p = __builtin_fma(r, p, 1.0 / 40320.0);
    p = __builtin_fma(r, p, 1.0 / 5040.0);
    p = __builtin_fma(r, p, 1.0 / 720.0);
    p = __builtin_fma(r, p, 1.0 / 120.0);
    p = __builtin_fma(r, p, 1.0 / 24.0);
    p = __builtin_fma(r, p, 1.0 / 6.0);
    p = __builtin_fma(r, p, 0.5);
    double er = __builtin_fma(r, __builtin_fma(r, p, 1.0), 1.0);
    int k1 = k / 2, k2 = k - k1;
    double s1 = __longlong_as_double((long long)((uint64_t)(k1 + 1023) << 52));
    double s2 = __longlong_as_double((long long)((uint64_t)(k2 + 1023) << 52));
    return (er * s1) * s2;
}

__device__ __forceinline__ double log_d(double x)
{
    if (x != x || x < 0.0) return __longlong_as_double(0x7ff8000000000000LL);
    if (x == 0.0) return -__longlong_as_double(0x7ff0000000000000LL);
    if (x == __longlong_as_double(0x7ff0000000000000LL)) return x;
    if (x < 2.2250738585072014e-308)
        return log_pos(x * 18014398509481984.0) - 54.0 * kLn2Hi - 54.0 * kLn2Lo;
    return log_pos(x);
}

static __constant__ const double kPowRecip[64] = {SW2D_POW_RECIP};
static __constant__ const double kPowNlog[64] = {SW2D_POW_NLOG};
static __constant__ const double kPowExp2[64] = {SW2D_POW_EXP2};

// pow(x, 1.3333330) for 2^-700 <= x <= 2^500 (table-driven, see the top)
__device__ __forceinline__ double pow1333_core(double x)
{
    const uint64_t b = (uint64_t)__double_as_longlong(x);
    const int e = (int)(b >> 52) - 1023;
    const int i = (int)((b >> 46) & 63);
    const double m = __longlong_as_double((long long)((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL));
    const double r = __builtin_fma(m, kPowRecip[i], -1.0);
    double q = 1.0 / 7.0;
    q = __builtin_fma(r, q, -1.0 / 6.0);
    q = __builtin_fma(r, q, 1.0 / 5.0);
    q = __builtin_fma(r, q, -0.25);
    q = __builtin_fma(r, q, 1.0 / 3.0);
    q = __builtin_fma(r, q, -0.5);
    const double lp = __builtin_fma(r * r, q, r);
    const double de = (double)e;
    const double lx = __builtin_fma(de, kLn2Hi, __builtin_fma(de, kLn2Lo, kPowNlog[i] + lp));
    const double t = kManningExp * lx;
    const double kd = __builtin_rint(t * (64.0 * kInvLn2));
    const int k = (int)kd;
    double rr = __builtin_fma(-kd, kLn2Hi / 64.0, t);
    rr = __builtin_fma(-kd, kLn2Lo / 64.0, rr);
    double p = 1.0 / 720.0;
    p = __builtin_fma(rr, p, 1.0 / 120.0);
    p = __builtin_fma(rr, p, 1.0 / 24.0);
    p = __builtin_fma(rr, p, 1.0 / 6.0);
    p = __builtin_fma(rr, p, 0.5);
    const double P = __builtin_fma(rr * rr, p, rr);
    const double T = kPowExp2[k & 63];
    return __builtin_ldexp(__builtin_fma(T, P, T), k >> 6);
}

// pow(x, 1.3333330) for FP64
__device__ __forceinline__ double pow1333(double x)
{
    if (x >= 0x1p-700 && x <= 0x1p500) return pow1333_core(x);
    if (x != x) return x;
    if (!(x > 0.0)) return x == 0.0 ? 0.0 : __longlong_as_double(0x7ff8000000000000LL);
    if (x == __longlong_as_double(0x7ff0000000000000LL)) return x;
    return exp_d(kManningExp * log_d(x));
}

// pow(x, 1.3333330) for FP32 (x > 0 on every path that uses the result)
__device__ __forceinline__ float pow1333(float x)
{
    return __builtin_amdgcn_exp2f((float)kManningExp * __builtin_amdgcn_logf(x));
}

__device__ __forceinline__ float exp_f(float t)
{
    return __builtin_amdgcn_exp2f(t * 1.44269504f);
}

// Arithmetic of the state type.  FP64: the IEEE operations, exactly as the
// reference writes them.  FP32 (precision form): hardware reciprocal and
// square root (about 1 ulp), division by a constant as a multiplication.
template <typename T> struct num;
template <> struct num<double> {
    __device__ static __forceinline__ double sqrt_(double x) { return __builtin_sqrt(x); }
    __device__ static __forceinline__ double abs_(double x) { return __builtin_fabs(x); }
    __device__ static __forceinline__ double max_(double a, double b) { return __builtin_fmax(a, b); }
    __device__ static __forceinline__ double exp_(double x) { return exp_d(x); }
    __device__ static __forceinline__ double div_(double a, double b) { return a / b; }
    // x / |x| for nonzero finite x (exactly +-1)
    __device__ static __forceinline__ double sign_(double x) { return __builtin_copysign(1.0, x); }
    // a / b where rb = 1 / b is precomputed (FP64 still divides)
    __device__ static __forceinline__ double divc_(double a, double b, double rb) { return a / b; }
};
template <> struct num<float> {
    __device__ static __forceinline__ float sqrt_(float x) { return __builtin_amdgcn_sqrtf(x); }
    __device__ static __forceinline__ float abs_(float x) { return __builtin_fabsf(x); }
    __device__ static __forceinline__ float max_(float a, float b) { return __builtin_fmaxf(a, b); }
    __device__ static __forceinline__ float exp_(float x) { return exp_f(x); }
    __device__ static __forceinline__ float div_(float a, float b) { return a * __builtin_amdgcn_rcpf(b); }
    __device__ static __forceinline__ float sign_(float x) { return __builtin_copysignf(1.0f, x); }
    __device__ static __forceinline__ float divc_(float a, float b, float rb) { return a * rb; }
};

// stream_flow's interpolation (SW2D-EFDC.cu:532-537, 546-551); reads past the
// end of the series give 0.  Exactly at t = k*dtser it divides 0 by 0, as the
// reference does.
__device__ __forceinline__ double series_at(const double *v, int n, int i)
{
    return (i >= 0 && i < n) ? v[i] : 0.0;
}
__device__ __forceinline__ double series_value(const double *v, int n, double dtser, double t)
{
    if (t <= 1.00) return series_at(v, n, 0) * t;
    int it = (int)(t / dtser);
    double a = series_at(v, n, it), b = series_at(v, n, it + 1);
    return a + (b - a) / (dtser * (t - dtser * (it)));
}
// point-source table Q[row * nst + i]
__device__ __forceinline__ double source_value(const double *Q, int nq, int nst, int i, double dtoq,
                                               double t)
{
    if (t <= 1.00) return Q[i] * t;
    int it = (int)(t / dtoq);
    double a = (it < nq) ? Q[(long)it * nst + i] : 0.0;
    double b = (it + 1 < nq) ? Q[(long)(it + 1) * nst + i] : 0.0;
    return a + (b - a) / (dtoq * (t - dtoq * (it)));
}

// Cross-lane moves along the wavefront with DPP (GFX9 wave_shr:1 / wave_shl:1).
// from_west: lane i receives lane i-1 (lane 0 receives 0); from_east: lane i
// receives lane i+1 (lane 63 receives 0).  All lanes must be active.  The
// bound-control bit makes the lane without a source read 0 itself, so the
// move needs no zeroed destination first (one instruction, foldable into the
// instruction that consumes it).
__device__ __forceinline__ int dpp_from_west(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, 0x138, 0xf, 0xf, true);
}
__device__ __forceinline__ int dpp_from_east(int v)
{
    return __builtin_amdgcn_update_dpp(0, v, 0x130, 0xf, 0xf, true);
}
__device__ __forceinline__ double from_west(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_from_west((int)(b & 0xffffffffLL)), hi = dpp_from_west((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ double from_east(double v)
{
    const long long b = __double_as_longlong(v);
    const int lo = dpp_from_east((int)(b & 0xffffffffLL)), hi = dpp_from_east((int)(b >> 32));
    return __longlong_as_double(((long long)hi << 32) | (long long)(unsigned)lo);
}
__device__ __forceinline__ float from_west(float v)
{
    return __int_as_float(dpp_from_west(__float_as_int(v)));
}
__device__ __forceinline__ float from_east(float v)
{
    return __int_as_float(dpp_from_east(__float_as_int(v)));
}

}  // namespace sw2d
