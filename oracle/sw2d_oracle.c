/*
 * sw2d_oracle.c -- CPU FP64 restatement of the SW2D overland-flow time step
 * of LabHidro/SW2D-EFDC (file SW2D-EFDC.cu).
 *
 * TEST INFRASTRUCTURE, NOT PRODUCT CODE.  Only tests/, the smoke check and the
 * cpu_baseline leg of bench.py load this library, and only as the checker of
 * the HIP library (or as the timed CPU baseline).  Nothing shipped links it.
 *
 * Layout follows the reference literally: one flat row-major array per field,
 * index c = r*C + j, r = 0 the north memory row, N = R*C + C/2 cells
 * (SW2D-EFDC.cu:3997) including the half "pad" row r = R.  Two documented
 * deviations from the reference, which reads memory it never initialised:
 *   - the pad row is inactive (inf = 0, h = 0);
 *   - every index past N reads 0 (a zeroed guard tail).
 * Both give the reference's results whenever that memory holds zeros
 * (pad values only ever reach real cells multiplied by a zero velocity).
 *
 * Every phase of one reference loop iteration (SW2D-EFDC.cu:4798-4877) is a
 * separate loop over cells, in the reference's order, with every expression
 * written in the reference's evaluation order; build with
 * -ffp-contract=off so that no multiply-add is fused.
 *
 * Extensions the reference does not have (adaptive CFL step, volume ledger,
 * outlet weir + hydrograph, lake-inflow accumulation, per-cell Manning n,
 * moving Gaussian storm) are hooks that stay off in reference mode; their
 * semantics are stated where they are implemented.
 *
 * Each function is pinned by tests/test_oracle.py (closed forms, invariants,
 * hand-computed branch values and the example case files of the reference).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

/* the three 64-entry tables of pow(x, 1.3333330) (plain data, correctly
 * rounded values; checked entry by entry in tests/test_oracle.py) */
#include "../sw2d-efdc_amd/csrc/sw2d_pow_tables.h"

#ifdef _OPENMP
#include <omp.h>
#endif

/* ------------------------------------------------------------------------
 * Project math: pow(x, 1.3333330) and exp(x) from correctly rounded operations
 * only (+ - * /, fma, floor, rint, ldexp, exponent bit manipulation, and
 * three tables of correctly rounded constants).  The HIP library evaluates
 * the identical sequence, so both sides round identically.
 *
 * Method: argument reduction x = 2^e * m, m in (sqrt(1/2), sqrt(2)];
 * log(m) = 2 atanh(s), s = (m-1)/(m+1), written as f - (hf - s(hf + R)) with
 * hf = f^2/2 and R(z) = sum_k 2 z^k/(2k+1), z = s^2 (the atanh series
 * arrangement of Sun's fdlibm e_log.c, with plain Taylor coefficients);
 * exp(t) = 2^k exp(r), r = t - k ln2 (Cody-Waite two-constant reduction),
 * exp(r) by its Taylor series to r^13.
 * ------------------------------------------------------------------------ */
#define ORC_LN2_HI 6.93147180369123816490e-01 /* ln2 high part, 21 low zero bits */
#define ORC_LN2_LO 1.90821492927058770002e-10 /* ln2 - ORC_LN2_HI */
#define ORC_INV_LN2 1.44269504088896338700e+00
#define ORC_SQRT2 1.41421356237309514547e+00
#define ORC_MANNING_EXP 1.3333330 /* SW2D-EFDC.cu:194, 311 */

static inline double orc_bits2d(uint64_t b) { double d; memcpy(&d, &b, 8); return d; }
static inline uint64_t orc_d2bits(double d) { uint64_t b; memcpy(&b, &d, 8); return b; }

/* natural log of a positive, finite, normal x */
static double orc_log_pos(double x)
{
    uint64_t b = orc_d2bits(x);
    int e = (int)((b >> 52) & 0x7ff) - 1023;
    double m = orc_bits2d((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    if (m > ORC_SQRT2) { m = m * 0.5; e = e + 1; }
    double f = m - 1.0;
    double s = f / (2.0 + f);
    double z = s * s;
    /* R(z) = sum_{k=1..12} 2 z^k / (2k+1), Horner with correctly rounded fma */
    double R = 2.0 / 25.0;
    R = fma(z, R, 2.0 / 23.0);
    R = fma(z, R, 2.0 / 21.0);
    R = fma(z, R, 2.0 / 19.0);
    R = fma(z, R, 2.0 / 17.0);
    R = fma(z, R, 2.0 / 15.0);
    R = fma(z, R, 2.0 / 13.0);
    R = fma(z, R, 2.0 / 11.0);
    R = fma(z, R, 2.0 / 9.0);
    R = fma(z, R, 2.0 / 7.0);
    R = fma(z, R, 2.0 / 5.0);
    R = fma(z, R, 2.0 / 3.0);
    R = z * R;
    double hf = 0.5 * f * f;
    double lm = f - (hf - s * (hf + R));
    double de = (double)e;
    return de * ORC_LN2_HI + (de * ORC_LN2_LO + lm);
}

double orc_exp(double t)
{
    if (t != t) return t;
    if (t > 709.0) return INFINITY;
    if (t < -708.0) return 0.0;
    double kd = floor(t * ORC_INV_LN2 + 0.5);
    int k = (int)kd;
    double r = (t - kd * ORC_LN2_HI) - kd * ORC_LN2_LO;
    /* Taylor to r^13, Horner with correctly rounded fma */
    double p = 1.0 / 6227020800.0;       /* 1/13! */
    p = fma(r, p, 1.0 / 479001600.0);    /* 1/12! */
    p = fma(r, p, 1.0 / 39916800.0);     /* 1/11! */
    p = fma(r, p, 1.0 / 3628800.0);      /* 1/10! */
    p = fma(r, p, 1.0 / 362880.0);       /* 1/9!  */
    p = fma(r, p, 1.0 / 40320.0);        /* 1/8!  */
    p = fma(r, p, 1.0 / 5040.0);         /* 1/7!  */
    p = fma(r, p, 1.0 / 720.0);          /* 1/6!  */
    p = fma(r, p, 1.0 / 120.0);          /* 1/5!  */
    p = fma(r, p, 1.0 / 24.0);           /* 1/4!  */
    p = fma(r, p, 1.0 / 6.0);            /* 1/3!  */
    p = fma(r, p, 0.5);                  /* 1/2!  */
    double er = fma(r, fma(r, p, 1.0), 1.0);
    /* 2^k split in two factors so that k down to -1022-52 stays exact */
    int k1 = k / 2, k2 = k - k1;
    double s1 = orc_bits2d((uint64_t)(k1 + 1023) << 52);
    double s2 = orc_bits2d((uint64_t)(k2 + 1023) << 52);
    return (er * s1) * s2;
}

double orc_log(double x)
{
    if (x != x || x < 0.0) return NAN;
    if (x == 0.0) return -INFINITY;
    if (x == INFINITY) return INFINITY;
    if (x < 2.2250738585072014e-308) return orc_log_pos(x * 18014398509481984.0) - 54.0 * ORC_LN2_HI - 54.0 * ORC_LN2_LO;
    return orc_log_pos(x);
}

static const double ORC_POW_RECIP[64] = {SW2D_POW_RECIP};
static const double ORC_POW_NLOG[64] = {SW2D_POW_NLOG};
static const double ORC_POW_EXP2[64] = {SW2D_POW_EXP2};

/* pow(x, 1.3333330) for 2^-700 <= x <= 2^500, table-driven without division:
 * x = 2^e m, m in [1, 2); i = top 6 fraction bits; r = m RECIP[i] - 1 (one
 * fma, |r| < 2^-7); ln x = e ln2 + NLOG[i] + log1p(r), log1p by Taylor to
 * r^7; t = 1.3333330 ln x; k = rint(64 t / ln2); r' = t - k ln2/64
 * (two-constant ln2); exp(t) = 2^(k>>6) EXP2[k&63] (1 + P(r')), P = Taylor of
 * exp(r') - 1 to r'^6.  Horner with correctly rounded fma. */
static double orc_pow1333_core(double x)
{
    uint64_t b = orc_d2bits(x);
    int e = (int)(b >> 52) - 1023;
    int i = (int)((b >> 46) & 63);
    double m = orc_bits2d((b & 0x000fffffffffffffULL) | 0x3ff0000000000000ULL);
    double r = fma(m, ORC_POW_RECIP[i], -1.0);
    double q = 1.0 / 7.0;
    q = fma(r, q, -1.0 / 6.0);
    q = fma(r, q, 1.0 / 5.0);
    q = fma(r, q, -0.25);
    q = fma(r, q, 1.0 / 3.0);
    q = fma(r, q, -0.5);
    double lp = fma(r * r, q, r);
    double de = (double)e;
    double lx = fma(de, ORC_LN2_HI, fma(de, ORC_LN2_LO, ORC_POW_NLOG[i] + lp));
    double t = ORC_MANNING_EXP * lx;
    double kd = rint(t * (64.0 * ORC_INV_LN2));
    int k = (int)kd;
    double rr = fma(-kd, ORC_LN2_HI / 64.0, t);
    rr = fma(-kd, ORC_LN2_LO / 64.0, rr);
    double p = 1.0 / 720.0;
    p = fma(rr, p, 1.0 / 120.0);
    p = fma(rr, p, 1.0 / 24.0);
    p = fma(rr, p, 1.0 / 6.0);
    p = fma(rr, p, 0.5);
    double P = fma(rr * rr, p, rr);
    double T = ORC_POW_EXP2[k & 63];
    return ldexp(fma(T, P, T), k >> 6);
}

/* pow(x, 1.3333330), the Manning depth exponent of SW2D-EFDC.cu:194, 311:
 * the table-driven form where nothing under- or overflows, else exp(c log x)
 * from the general functions above */
double orc_pow1333(double x)
{
    if (x >= 0x1p-700 && x <= 0x1p500) return orc_pow1333_core(x);
    if (x != x) return x;
    if (!(x > 0.0)) return x == 0.0 ? 0.0 : NAN;
    if (x == INFINITY) return INFINITY;
    return orc_exp(ORC_MANNING_EXP * orc_log(x));
}

/* ------------------------------------------------------------------------
 * Host forcing: SW2D-EFDC.cu:525-554 (stream_flow) and 588-600
 * (evaporation_load).  The interpolation divides by
 * dtrain*(t - dtrain*it), as the reference does, which gives 0/0 = NaN or
 * +-inf exactly at t = k*dtrain (kept on purpose).  Reads past the end of a
 * series return 0 (the reference reads unallocated memory there).
 * ------------------------------------------------------------------------ */
static inline double orc_series_at(const double *v, int n, int i)
{
    return (i >= 0 && i < n) ? v[i] : 0.0;
}

/* value of a series at time t, in the series' own units */
double orc_series(const double *v, int n, double dtser, double t)
{
    if (t <= 1.00) return orc_series_at(v, n, 0) * t;
    int it = (int)(t / dtser);
    double a = orc_series_at(v, n, it), b = orc_series_at(v, n, it + 1);
    return a + (b - a) / (dtser * (t - dtser * (it)));
}

/* rain rate rr in m/s at time t (SW2D-EFDC.cu:546-553) */
double orc_rain_rate(const double *rain, int n, double dtrain, double t)
{
    double rr = orc_series(rain, n, dtrain, t);
    return rr / (dtrain * 1000.0);
}

/* flat cell index of a point source (SW2D-EFDC.cu:541-544); xcoor = xll/res,
 * brx = x/res, ycoor = yll/res, bry = y/res as in the reference */
int orc_source_index(int rows, int cols, double xcoor, double ycoor, double brx, double bry)
{
    int qinx = (int)round(fabs(xcoor - brx));
    int qiny = rows - (int)round(fabs(ycoor - bry));
    return (qiny + 1) * cols - (cols - (qinx + 1));
}

/* evaporation series in mm per interval (SW2D-EFDC.cu:556-586) */
void orc_evaporation(double albedo, const double *T, const double *Rg, int n, double dtime, double *Evapo)
{
    for (int i = 0; i < n; i++) {
        double Rs = (1 - albedo) * Rg[i];
        double pw = 999.85308 + 6.32693 * (0.01) * T[i] - 8.523829 * (0.001) * T[i] * T[i] +
                    6.943248 * (0.00001) * T[i] * T[i] * T[i] -
                    3.821216 * (0.0000001) * T[i] * T[i] * T[i] * T[i];
        double lv = (((T[i] + 273.15) / ((T[i] + 273.15) - 33.91)) *
                     ((T[i] + 273.15) / ((T[i] + 273.15) - 33.91))) * 1.91846 * 1000000;
        double E = (Rs / (lv * pw)) * dtime * 1000.0;
        if (E < 0) E = 0;
        Evapo[i] = E;
    }
}

/* ------------------------------------------------------------------------
 * State
 * ------------------------------------------------------------------------ */
enum { ORC_F_H = 0, ORC_F_UM, ORC_F_VN, ORC_F_UU, ORC_F_VV, ORC_F_UUA, ORC_F_VVA,
       ORC_F_Z, ORC_F_HM, ORC_F_HN, ORC_F_UU1, ORC_F_VV1, ORC_F_INF, ORC_F_COUNT };

typedef struct orc {
    int R, C, N, NA;          /* rows, cols, reference N, allocated length */
    double dx, dy, g, rn, th, dt, nodata;
    double INT, INF, LWL, EV_WL_min;
    double time;
    long mstep;
    /* flat fields, each NA long, index 0..N-1 real + pad, N..NA-1 zero guard */
    double *z, *h, *ho, *hm, *hn, *um, *umo, *uu1, *vv1, *uua, *vva, *vn, *vno, *uu, *vv, *ql;
    int *inf;
    /* forcing series */
    double *rain; int nrain; double dtrain;
    double *Q; int nq, nst; double dtoq; int *src;  /* Q[row*nst + i], src[i] flat index */
    double *evapo; int nevapo; int evap_on;
    /* extension: per-cell Manning n (NULL = scalar rn) */
    double *nmap;
    /* extension: moving Gaussian storm replacing the rain series */
    int storm_on; double st_peak, st_x0, st_y0, st_vx, st_vy, st_sigma, st_t0, st_t1;
    /* extension: free-outflow outlet cell on a domain edge */
    int out_r, out_c, out_edge;  /* edge: 0 none, 1 south, 2 north, 3 west, 4 east */
    /* extension: adaptive CFL step */
    int adaptive; double courant, dt2_max, dt2_next;
    /* ledgers (extensions) */
    double v0, v_src, v_clip, v_out, q_out_last, q_out_peak, t_peak;
    double *hydro; long nhydro, hydro_cap;   /* per-iteration (t, Q_out) pairs */
    /* extension: lake-inflow accumulation over coupling records */
    int nbc; int *bc_cell; int *bc_kind; double *bc_acc;  /* kind 0 S,1 N,2 E,3 W */
    int nthreads;
} orc_t;

static double *orc_zalloc(int n) { return (double *)calloc((size_t)n, sizeof(double)); }

static double orc_volume(const orc_t *s)
{
    double v = 0.0;
    for (int c = 0; c < s->R * s->C; c++)
        if (s->inf[c]) v += s->h[c];
    return v * s->dx * s->dy;
}

/* initial masks and face depths: init_inf (SW2D-EFDC.cu:29-49) and the hm/hn
 * part of initiald (51-91); infx/infy are not kept because for every face
 * that passes the row/column guard of flux they equal |inf - inf'| and
 * are therefore implied by the two inf tests. */
static void orc_init_fields(orc_t *s)
{
    const int R = s->R, C = s->C, N = s->N;
    for (int c = 0; c < N; c++) {
        int ini = c / C;
        s->ho[c] = s->h[c];
        s->inf[c] = (ini < R) ? (s->z[c] != s->nodata) : 0;  /* pad row inactive */
    }
    for (int c = 0; c < N; c++) {
        int inj = c % C, ini = c / C;
        if (ini == 0) s->hm[c] = s->h[c];
        else if (ini == R) s->hm[c] = s->h[c - C];
        else s->hm[c] = 0.50 * (s->h[c] + s->h[c - C]);
        if (inj == 0) s->hn[c] = s->h[c];
        else s->hn[c] = 0.50 * (s->h[c] + s->h[c - 1]);
    }
}

void orc_cfl_update(orc_t *s);
static double orc_cfl_min(const orc_t *s);

orc_t *orc_create(int R, int C, double cellsize, double nodata, const double *z, const double *h0,
                  double g, double rn, double th, double dt, double INT, double INF, double LWL,
                  double EV_WL_min, double time0)
{
    orc_t *s = (orc_t *)calloc(1, sizeof(orc_t));
    s->R = R; s->C = C; s->N = R * C + C / 2; s->NA = s->N + 2 * C + 4;
    s->dx = cellsize; s->dy = cellsize; s->g = g; s->rn = rn; s->th = th; s->dt = dt;
    s->nodata = nodata; s->INT = INT; s->INF = INF; s->LWL = LWL; s->EV_WL_min = EV_WL_min;
    s->time = time0; s->mstep = 0;
    double **f[] = {&s->z, &s->h, &s->ho, &s->hm, &s->hn, &s->um, &s->umo, &s->uu1, &s->vv1,
                    &s->uua, &s->vva, &s->vn, &s->vno, &s->uu, &s->vv, &s->ql};
    for (unsigned i = 0; i < sizeof(f) / sizeof(f[0]); i++) *f[i] = orc_zalloc(s->NA);
    s->inf = (int *)calloc((size_t)s->NA, sizeof(int));
    memcpy(s->z, z, sizeof(double) * (size_t)R * C);
    memcpy(s->h, h0, sizeof(double) * (size_t)R * C);
    orc_init_fields(s);
    /* inactive cells carry no water in any later state (continuity zeroes them);
     * the initial depth is taken as given, as the reference does */
    s->out_r = -1; s->out_c = -1; s->out_edge = 0;
    s->dt2_max = 2.00 * dt; s->courant = 0.5;
    s->v0 = orc_volume(s);
    s->nthreads = 1;
    return s;
}

void orc_destroy(orc_t *s)
{
    if (!s) return;
    double *f[] = {s->z, s->h, s->ho, s->hm, s->hn, s->um, s->umo, s->uu1, s->vv1, s->uua,
                   s->vva, s->vn, s->vno, s->uu, s->vv, s->ql, s->rain, s->Q, s->evapo, s->nmap,
                   s->hydro, s->bc_acc};
    for (unsigned i = 0; i < sizeof(f) / sizeof(f[0]); i++) free(f[i]);
    free(s->inf); free(s->src); free(s->bc_cell); free(s->bc_kind);
    free(s);
}

void orc_set_threads(orc_t *s, int n) { s->nthreads = n > 0 ? n : 1; }

void orc_set_rain(orc_t *s, const double *rain, int n, double dtrain)
{
    free(s->rain);
    s->rain = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (n > 0) memcpy(s->rain, rain, sizeof(double) * (size_t)n);
    s->nrain = n; s->dtrain = dtrain;
}

/* sources: flat indices (already computed with orc_source_index) and the
 * Q_source_sink table, nq rows of nst values (m3/s), spacing dtoq */
int orc_set_sources(orc_t *s, const int *src, int nst, const double *Q, int nq, double dtoq)
{
    free(s->src); free(s->Q);
    s->src = (int *)malloc(sizeof(int) * (size_t)(nst > 0 ? nst : 1));
    s->Q = (double *)malloc(sizeof(double) * (size_t)(nq * nst > 0 ? nq * nst : 1));
    for (int i = 0; i < nst; i++) {
        if (src[i] < 0 || src[i] >= s->R * s->C) return -1;
        s->src[i] = src[i];
    }
    if (nq * nst > 0) memcpy(s->Q, Q, sizeof(double) * (size_t)nq * nst);
    s->nst = nst; s->nq = nq; s->dtoq = dtoq;
    return 0;
}

void orc_set_evapo(orc_t *s, const double *evapo, int n)
{
    free(s->evapo);
    s->evapo = (double *)malloc(sizeof(double) * (size_t)(n > 0 ? n : 1));
    if (n > 0) memcpy(s->evapo, evapo, sizeof(double) * (size_t)n);
    s->nevapo = n; s->evap_on = 1;
}

/* per-cell Manning n, R*C values; the face coefficient is 0.5*(n_w + n_e) */
void orc_set_manning_map(orc_t *s, const double *nmap)
{
    free(s->nmap);
    s->nmap = orc_zalloc(s->NA);
    memcpy(s->nmap, nmap, sizeof(double) * (size_t)s->R * s->C);
}

/* moving Gaussian storm: rate(x,y,t) = peak * (gx * gy) for t0 <= t < t1,
 * gx = exp(-(x - (x0 + vx t))^2 / (2 sigma^2)), gy likewise in y, with
 * x = (j + 0.5) dx, y = (r + 0.5) dy measured from the north-west corner;
 * peak in m/s */
void orc_set_storm(orc_t *s, double peak, double x0, double y0, double vx, double vy,
                   double sigma, double t0, double t1)
{
    s->storm_on = 1; s->st_peak = peak; s->st_x0 = x0; s->st_y0 = y0; s->st_vx = vx;
    s->st_vy = vy; s->st_sigma = sigma; s->st_t0 = t0; s->st_t1 = t1;
}

static inline double orc_storm_factor(double d, double sigma)
{
    return orc_exp(-(d * d) / (2.0 * sigma * sigma));
}

/* free-outflow outlet: a cell on a domain edge whose edge face passes the
 * critical-flow weir discharge q = 0.5440*h*sqrt(g*h) (the formula of
 * SW2D-EFDC.cu:131) computed from the depth at the start of the iteration */
int orc_set_outlet(orc_t *s, int r, int c)
{
    if (r < 0 || r >= s->R || c < 0 || c >= s->C) return -1;
    int edge = 0;
    if (r == s->R - 1) edge = 1;
    else if (r == 0) edge = 2;
    else if (c == 0) edge = 3;
    else if (c == s->C - 1) edge = 4;
    if (!edge) return -1;
    s->out_r = r; s->out_c = c; s->out_edge = edge;
    return 0;
}

void orc_set_hydrograph(orc_t *s, long cap)
{
    free(s->hydro);
    s->hydro = orc_zalloc((int)(2 * cap));
    s->hydro_cap = cap; s->nhydro = 0;
}

/* adaptive step: iteration k+1 has length dt2 = min(dt2_max, courant * m)
 * with m the CFL minimum (orc_cfl_min) of the state at the START of
 * iteration k (a one-iteration lag, so that a device kernel can form it
 * from the rows it streams in); iterations 0 and 1 both use m of the
 * initial state */
void orc_set_adaptive(orc_t *s, double courant, double dt2_max)
{
    s->adaptive = 1; s->courant = courant; s->dt2_max = dt2_max;
    orc_cfl_update(s);
}

/* lake-inflow records: flat cell + kind (0 "South" p-C, 1 "North" p+C,
 * 2 "East" p-1, 3 "West" p+1 in couple_BC_efdc's naming) */
void orc_set_bc(orc_t *s, const int *cell, const int *kind, int n)
{
    free(s->bc_cell); free(s->bc_kind); free(s->bc_acc);
    s->nbc = n;
    s->bc_cell = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    s->bc_kind = (int *)malloc(sizeof(int) * (size_t)(n > 0 ? n : 1));
    s->bc_acc = orc_zalloc(n > 0 ? n : 1);
    for (int i = 0; i < n; i++) { s->bc_cell[i] = cell[i]; s->bc_kind[i] = kind[i]; }
}

/* inflow rate of one coupling record (the volumeFlow_BC == 1 formulas of
 * SW2D-EFDC.cu:764-821) from depth h, uua and vva at the record's cell */
double orc_bc_rate(int kind, double h, double uua, double vva, double dx, double dy)
{
    switch (kind) {
    case 0: return (uua < 0) ? -uua * h * dx : 0.00;
    case 1: return (uua > 0) ? uua * h * dx : 0.00;
    case 2: return (vva < 0) ? -vva * h * dy : 0.00;
    default: return (vva > 0) ? vva * h * dy : 0.00;
    }
}

/* ------------------------------------------------------------------------
 * Phases of one loop iteration
 * ------------------------------------------------------------------------ */

/* flux, SW2D-EFDC.cu:97-353 */
static void orc_flux(orc_t *s, double dt2)
{
    const int R = s->R, C = s->C, N = s->N;
    const double th = s->th, gg = s->g, dx = s->dx, dy = s->dy;
    const double *h = s->h, *z = s->z, *hm = s->hm, *hn = s->hn, *uu1 = s->uu1, *vv1 = s->vv1;
    const double *umo = s->umo, *vno = s->vno, *uua = s->uua, *vva = s->vva;
    const int *inf = s->inf;
    double *um = s->um, *vn = s->vn;
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
    for (int c = 0; c < N; c++) {
        int inj = c % C, ini = c / C;
        /* x direction: row faces, um at the north face of c */
        double out = 0;
        if (ini > 0 && ini < R && inf[c] != 0 && inf[c - C] != 0 &&
            ((h[c - C] > th) || (h[c] > th))) {
            double hhe = h[c] + z[c], hhw = h[c - C] + z[c - C];
            double hhep = h[c] - th, hhwp = h[c - C] - th;
            if (hhe < z[c - C]) {
                out = (h[c - C] > th) ? 0.5440 * h[c - C] * sqrt(gg * h[c - C]) : 0;
            } else if (hhw < z[c]) {
                out = (h[c] > th) ? -0.544 * h[c] * sqrt(gg * h[c]) : 0;
            } else if (hhep * hhwp < 0) {
                if ((h[c] > 0) || (h[c - C] > 0)) {
                    double hhan = hhep - hhwp;
                    double sgnm = hhan / fabs(hhan);
                    double hh3 = fmax((h[c] + z[c]), (h[c - C] + z[c - C])) - fmax(z[c], z[c - C]);
                    out = -sgnm * 0.350 * hh3 * sqrt(2.00 * gg * hh3);
                } else out = 0;
            } else {
                double u13 = gg * hm[c] * (dt2 / dx) * (h[c] + z[c] - h[c - C] - z[c - C]);
                double ur = 0.50 * (uu1[c + C] + uu1[c]);
                double ul = 0.50 * (uu1[c] + uu1[c - C]);
                double umr = ur * (umo[c + C] + umo[c]) * 0.50 + fabs(ur) * (umo[c] - umo[c + C]) * 0.50;
                double uml = ul * (umo[c] + umo[c - C]) * 0.50 + fabs(ul) * (umo[c - C] - umo[c]) * 0.50;
                double u11 = (dt2 / dx) * (umr - uml);
                double u12;
                if (inj == 0) u12 = 0;
                else {
                    double vu = 0.50 * (vv1[c + 1] + vv1[c - C + 1]);
                    double vd = 0.50 * (vv1[c] + vv1[c - C]);
                    double umu = vu * (umo[c + 1] + umo[c]) * 0.50 + fabs(vu) * (umo[c] - umo[c + 1]) * 0.50;
                    double umd = vd * (umo[c - 1] + umo[c]) * 0.50 + fabs(vd) * (umo[c - 1] - umo[c]) * 0.50;
                    u12 = (dt2 / dy) * (umu - umd);
                }
                double nf = s->nmap ? 0.5 * (s->nmap[c - C] + s->nmap[c]) : s->rn;
                double sqx = sqrt(uu1[c] * uu1[c] + vva[c] * vva[c]);
                double expx = orc_pow1333(hm[c]);
                double ram = gg * nf * nf * sqx / expx;
                if (hm[c] <= th) ram = 0.00;
                out = ((1.00 - dt2 * ram * 0.50) * umo[c] + (-u11 - u12 - u13)) / (1.00 + dt2 * ram * 0.50);
            }
        }
        um[c] = out;
        /* y direction: column faces, vn at the west face of c */
        out = 0;
        if (inj > 0 && inj < C && inf[c] != 0 && inf[c - 1] != 0 &&
            ((h[c - 1] > th) || (h[c] > th))) {
            double hhn = h[c] + z[c], hhs = h[c - 1] + z[c - 1];
            double hhnp = h[c] - th, hhsp = h[c - 1] - th;
            if (hhn < z[c - 1]) {
                out = (h[c - 1] > th) ? 0.5440 * h[c - 1] * sqrt(gg * h[c - 1]) : 0;
            } else if (hhs < z[c]) {
                out = (h[c] > th) ? -0.544 * h[c] * sqrt(gg * h[c]) : 0;
            } else if (hhnp * hhsp < 0) {
                if ((h[c] > 0) || (h[c - 1] > 0)) {
                    double hhan = hhnp - hhsp;
                    double sgnm = hhan / fabs(hhan);
                    double hh3 = fmax((h[c] + z[c]), (h[c - 1] + z[c - 1])) - fmax(z[c], z[c - 1]);
                    out = -sgnm * 0.350 * hh3 * sqrt(2.00 * gg * hh3);
                } else out = 0;
            } else {
                double v13 = gg * hn[c] * (dt2 / dy) * (h[c] + z[c] - h[c - 1] - z[c - 1]);
                double v11;
                if (ini == 0) v11 = 0;
                else if (ini == R) v11 = 0;
                else {
                    double a = 0.50 * (uu1[c + C] + uu1[c + C - 1]);
                    double b = 0.50 * (uu1[c] + uu1[c - 1]);
                    double vnr = a * (vno[c + C] + vno[c]) * 0.50 + fabs(a) * (vno[c] - vno[c + C]) * 0.50;
                    double vnl = b * (vno[c] + vno[c - C]) * 0.50 + fabs(b) * (vno[c - C] - vno[c]) * 0.50;
                    v11 = (dt2 / dx) * (vnr - vnl);
                }
                double p = 0.50 * (vv1[c + 1] + vv1[c]);
                double q = 0.50 * (vv1[c] + vv1[c - 1]);
                double vnu = p * (vno[c + 1] + vno[c]) * 0.50 + fabs(p) * (vno[c] - vno[c + 1]) * 0.50;
                double vnd = q * (vno[c - 1] + vno[c]) * 0.50 + fabs(q) * (vno[c - 1] - vno[c]) * 0.50;
                double v12 = (dt2 / dy) * (vnu - vnd);
                double nf = s->nmap ? 0.5 * (s->nmap[c - 1] + s->nmap[c]) : s->rn;
                double sqy = sqrt(uua[c] * uua[c] + vv1[c] * vv1[c]);
                double expy = orc_pow1333(hn[c]);
                double ram = gg * nf * nf * sqy / expy;
                if (hn[c] <= th) ram = 0.00;
                out = ((1.00 - dt2 * ram * 0.50) * vno[c] + (-v11 - v12 - v13)) / (1.00 + dt2 * ram * 0.50);
            }
        }
        vn[c] = out;
    }
}

/* forcing at one half-step */
typedef struct { double rr, Ev, gx_t, gy_t; int storm; } orc_forcing_t;

static void orc_forcing(orc_t *s, double t, orc_forcing_t *fo)
{
    /* stream_flow, SW2D-EFDC.cu:525-554: dense ql written at each source */
    for (int i = 0; i < s->nst; i++) {
        double ql;
        if (t <= 1.00) ql = s->Q[0 * s->nst + i] * t;
        else {
            int it = (int)(t / s->dtoq);
            double a = (it < s->nq) ? s->Q[it * s->nst + i] : 0.0;
            double b = (it + 1 < s->nq) ? s->Q[(it + 1) * s->nst + i] : 0.0;
            ql = a + (b - a) / (s->dtoq * (t - s->dtoq * (it)));
        }
        ql = ql / (s->dx * s->dy);
        s->ql[s->src[i]] = ql;
    }
    fo->rr = (s->nrain > 0) ? orc_rain_rate(s->rain, s->nrain, s->dtrain, t) : 0.0;  /* no rain.dat series: no rain */
    fo->Ev = 0.0;
    if (s->evap_on) fo->Ev = orc_series(s->evapo, s->nevapo, s->dtrain, t) / (s->dtrain * 1000.0);
    fo->storm = s->storm_on && (t >= s->st_t0) && (t < s->st_t1);
    fo->gx_t = s->st_x0 + s->st_vx * t;
    fo->gy_t = s->st_y0 + s->st_vy * t;
}

static inline double orc_cell_rain(const orc_t *s, const orc_forcing_t *fo, int c)
{
    if (!s->storm_on) return fo->rr;
    if (!fo->storm) return 0.0;
    int r = c / s->C, j = c % s->C;
    double x = ((double)j + 0.5) * s->dx, y = ((double)r + 0.5) * s->dy;
    double gx = orc_storm_factor(x - fo->gx_t, s->st_sigma);
    double gy = orc_storm_factor(y - fo->gy_t, s->st_sigma);
    return s->st_peak * (gx * gy);
}

/* continuity, SW2D-EFDC.cu:602-664; the outlet replaces its (zero) edge
 * face by the weir discharge qo; ledger = 1 accumulates the volume terms */
static void orc_continuity(orc_t *s, double dt2, const orc_forcing_t *fo, double qo, int ledger)
{
    const int C = s->C, N = s->N;
    const double dx = s->dx, dy = s->dy;
    const double P_lake = 1.00 - s->LWL, P_land = 1.00 - (s->INT + s->INF);
    const int oc = (s->out_edge) ? s->out_r * C + s->out_c : -1;
    double src_sum = 0.0, clip_sum = 0.0;
#pragma omp parallel for schedule(static) num_threads(s->nthreads) reduction(+ : src_sum, clip_sum)
    for (int c = 0; c < N; c++) {
        double P, E;
        if (s->ho[c] > s->EV_WL_min) { P = P_lake; E = fo->Ev; }
        else { P = P_land; E = 0.000; }
        double rr = orc_cell_rain(s, fo, c);
        double ums = s->um[c + C], umn = s->um[c], vne = s->vn[c + 1], vnw = s->vn[c];
        if (c == oc) {
            if (s->out_edge == 1) ums = qo;
            else if (s->out_edge == 2) umn = -qo;
            else if (s->out_edge == 3) vnw = -qo;
            else vne = qo;
        }
        double hr = s->ho[c] - dt2 * ((ums - umn) / dx + (vne - vnw) / dy - s->ql[c] - rr * P + E);
        double hv = fmax(hr, 0.00);
        if (s->inf[c] == 0) hv = 0.00;
        s->h[c] = hv;
        if (ledger && s->inf[c]) {
            src_sum += dt2 * (s->ql[c] + rr * P - E);
            if (hr < 0.0) clip_sum += -hr;
        }
    }
    if (ledger) {
        s->v_src += src_sum * dx * dy;
        s->v_clip += clip_sum * dx * dy;
    }
}

/* treat_error, SW2D-EFDC.cu:695-725, written per face: a face is zeroed when
 * its flux leaves a dry active cell (h < th).  The reference writes the same
 * zeros from the cell side; the tests are sign-exclusive, so the result does
 * not depend on the order of the writes. */
static void orc_treat_error(orc_t *s)
{
    const int C = s->C, N = s->N, R = s->R;
    const double th = s->th;
    double *um = s->um, *vn = s->vn;
    const double *h = s->h;
    const int *inf = s->inf;
#define ORC_DRY(c) (((inf[c] == 1) || (inf[c] == 2)) && (h[c] < th))
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
    for (int f = 0; f < N + C; f++) {
        if (f >= C && (f - C) / C < R && f - C < N && ORC_DRY(f - C) && um[f] > 0) um[f] = 0;
        if (f < N && ORC_DRY(f) && um[f] < 0.0) um[f] = 0.00;
        if (f >= 1 && f - 1 < N && ORC_DRY(f - 1) && vn[f] > 0) vn[f] = 0;
        if (f < N && ORC_DRY(f) && vn[f] < 0) vn[f] = 0;
    }
#undef ORC_DRY
}

/* hm_hn 362-398, uu1_vv1 404-452, uu_vv 458-483, uua_vva 489-522, forward 670-693 */
static void orc_derived(orc_t *s)
{
    const int C = s->C, N = s->N, R = s->R;
    const double th = s->th;
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
    for (int c = 0; c < N; c++) {
        int inj = c % C, ini = c / C;
        if (inj > 0) s->hn[c] = 0.50 * (s->h[c] + s->h[c - 1]);
        if (ini > 0) s->hm[c] = 0.50 * (s->h[c] + s->h[c - C]);
        if (ini == 0) s->hm[c] = s->h[c];
        if (ini == R) s->hm[c] = s->h[c - C];
        if (inj == 0) s->hn[c] = s->h[c];
    }
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
    for (int c = 0; c < N; c++) {
        s->uu1[c] = (s->hm[c] >= th) ? s->um[c] / s->hm[c] : 0.00;
        s->vv1[c] = (s->hn[c] >= th) ? s->vn[c] / s->hn[c] : 0.0;
    }
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
    for (int c = 0; c < N; c++) {
        if (s->h[c] >= th) {
            s->uu[c] = (s->uu1[c + C] + s->uu1[c]) / 2.00;
            s->vv[c] = (s->vv1[c + 1] + s->vv1[c]) / 2.00;
        } else { s->uu[c] = 0.00; s->vv[c] = 0.00; }
    }
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
    for (int c = 0; c < N; c++) {
        int inj = c % C, ini = c / C;
        if (inj > 0) s->uua[c] = 0.250 * (s->uu1[c] + s->uu1[c + C] + s->uu1[c - 1] + s->uu1[c + C - 1]);
        if (ini > 0) s->vva[c] = 0.250 * (s->vv1[c] + s->vv1[c + 1] + s->vv1[c - C] + s->vv1[c - C + 1]);
        if (inj == 0) s->uua[c] = 0.50 * (s->uu1[c] + s->uu1[c + C]);
        if (ini == 0) s->vva[c] = 0.50 * (s->vv1[c] + s->vv1[c + 1]);
        if (ini == R) s->vva[c] = 0.50 * (s->vv1[c] + s->vv1[c + 1]);
    }
#pragma omp parallel for schedule(static) num_threads(s->nthreads)
    for (int c = 0; c < N; c++) {
        s->umo[c] = s->um[c]; s->vno[c] = s->vn[c]; s->ho[c] = s->h[c];
    }
}

/* minimum over interior faces joining two active cells whose face depth
 * exceeds th of dx/(|uu1| + sqrt(g*hm)) (row faces) and dy/(|vv1| +
 * sqrt(g*hn)) (column faces), on the derived fields of the current state */
static double orc_cfl_min(const orc_t *s)
{
    const int C = s->C, R = s->R;
    double m = INFINITY;
    /* a minimum does not depend on the order it is taken in */
#pragma omp parallel for schedule(static) num_threads(s->nthreads) reduction(min : m)
    for (int r = 0; r < R; r++)
        for (int j = 0; j < C; j++) {
            int c = r * C + j;
            if (r > 0 && s->inf[c] && s->inf[c - C] && s->hm[c] > s->th) {
                double w = s->dx / (fabs(s->uu1[c]) + sqrt(s->g * s->hm[c]));
                if (w < m) m = w;
            }
            if (j > 0 && s->inf[c] && s->inf[c - 1] && s->hn[c] > s->th) {
                double w = s->dy / (fabs(s->vv1[c]) + sqrt(s->g * s->hn[c]));
                if (w < m) m = w;
            }
        }
    return m;
}

static double orc_cfl_dt2(const orc_t *s, double m)
{
    double d = s->courant * m;
    return (d < s->dt2_max) ? d : s->dt2_max;
}

void orc_cfl_update(orc_t *s) { s->dt2_next = orc_cfl_dt2(s, orc_cfl_min(s)); }

/* one reference loop iteration, SW2D-EFDC.cu:4798-4877 */
void orc_iteration(orc_t *s)
{
    double dt, dt2;
    if (s->adaptive) { dt2 = s->dt2_next; dt = 0.5 * dt2; }
    else { dt = s->dt; dt2 = 2.00 * dt; }
    const int C = s->C;
    const double m_k = s->adaptive ? orc_cfl_min(s) : 0.0;

    /* lake-inflow accumulation over the state at the start of the iteration */
    for (int i = 0; i < s->nbc; i++) {
        int p = s->bc_cell[i];
        s->bc_acc[i] += orc_bc_rate(s->bc_kind[i], s->h[p], s->uua[p], s->vva[p], s->dx, s->dy) * dt2;
    }
    /* outlet weir discharge from the depth at the start of the iteration */
    double qo = 0.0;
    if (s->out_edge) {
        double ho = s->h[s->out_r * C + s->out_c];
        qo = (ho > s->th) ? 0.5440 * ho * sqrt(s->g * ho) : 0;
    }

    orc_flux(s, dt2);
    orc_forcing_t fo;
    orc_forcing(s, s->time, &fo);
    orc_continuity(s, dt2, &fo, qo, 0);
    if (s->out_edge) {  /* treat_error on the edge face: zero it if the outlet cell is dry */
        int oc = s->out_r * C + s->out_c;
        if (((s->inf[oc] == 1) || (s->inf[oc] == 2)) && (s->h[oc] < s->th)) qo = 0;
    }
    orc_treat_error(s);
    s->time = s->time + dt;
    s->mstep = s->mstep + 1;
    orc_forcing(s, s->time, &fo);
    orc_continuity(s, dt2, &fo, qo, 1);
    orc_derived(s);
    s->time = s->time + dt;
    s->mstep = s->mstep + 1;

    double Qout = qo * s->dy;
    s->v_out += Qout * dt2;
    s->q_out_last = Qout;
    if (Qout > s->q_out_peak) { s->q_out_peak = Qout; s->t_peak = s->time; }
    if (s->hydro && s->nhydro < s->hydro_cap) {
        s->hydro[2 * s->nhydro] = s->time;
        s->hydro[2 * s->nhydro + 1] = Qout;
        s->nhydro++;
    }
    if (s->adaptive) s->dt2_next = orc_cfl_dt2(s, m_k);
}

/* test aid: round the state (h, um, vn and their "old" copies) to binary32
 * and rebuild the derived fields, to measure what FP32 storage alone costs */
void orc_round_state_f32(orc_t *s)
{
    for (int c = 0; c < s->NA; c++) {
        s->h[c] = (double)(float)s->h[c];
        s->um[c] = (double)(float)s->um[c];
        s->vn[c] = (double)(float)s->vn[c];
    }
    orc_derived(s);
}

/* test aid: replace (h, um, vn) of the R*C real cells and recompute the
 * derived fields, as the end of an iteration does (hm_hn .. forward) */
void orc_set_state(orc_t *s, const double *h, const double *um, const double *vn)
{
    const size_t RC = (size_t)s->R * s->C;
    memcpy(s->h, h, sizeof(double) * RC);
    memcpy(s->um, um, sizeof(double) * RC);
    memcpy(s->vn, vn, sizeof(double) * RC);
    orc_derived(s);
}

void orc_step(orc_t *s, int n)
{
    for (int i = 0; i < n; i++) orc_iteration(s);
}

/* run while time + dt <= tmax (the reference loop condition, 4798), at most
 * max_iters iterations; returns the number of iterations run */
long orc_run_until(orc_t *s, double tmax, long max_iters)
{
    long k = 0;
    for (;;) {
        double dt = s->adaptive ? 0.5 * s->dt2_next : s->dt;
        if (!(s->time + dt <= tmax) || k >= max_iters) break;
        orc_iteration(s);
        k++;
    }
    return k;
}

double orc_time(const orc_t *s) { return s->time; }
long orc_mstep(const orc_t *s) { return s->mstep; }
double orc_dt2_next(const orc_t *s) { return s->adaptive ? s->dt2_next : 2.00 * s->dt; }

/* copy a field's R*C real cells in memory order */
int orc_get_field(const orc_t *s, int field, double *out)
{
    const double *src = NULL;
    switch (field) {
    case ORC_F_H: src = s->h; break;
    case ORC_F_UM: src = s->um; break;
    case ORC_F_VN: src = s->vn; break;
    case ORC_F_UU: src = s->uu; break;
    case ORC_F_VV: src = s->vv; break;
    case ORC_F_UUA: src = s->uua; break;
    case ORC_F_VVA: src = s->vva; break;
    case ORC_F_Z: src = s->z; break;
    case ORC_F_HM: src = s->hm; break;
    case ORC_F_HN: src = s->hn; break;
    case ORC_F_UU1: src = s->uu1; break;
    case ORC_F_VV1: src = s->vv1; break;
    case ORC_F_INF:
        for (int c = 0; c < s->R * s->C; c++) out[c] = (double)s->inf[c];
        return 0;
    default: return -1;
    }
    memcpy(out, src, sizeof(double) * (size_t)s->R * s->C);
    return 0;
}

/* the volume ledger: out[0] volume now, [1] initial volume, [2] source volume,
 * [3] clip volume, [4] outflow volume, [5] last outflow rate, [6] peak outflow,
 * [7] time of the peak */
void orc_ledger(const orc_t *s, double *out)
{
    out[0] = orc_volume(s); out[1] = s->v0; out[2] = s->v_src; out[3] = s->v_clip;
    out[4] = s->v_out; out[5] = s->q_out_last; out[6] = s->q_out_peak; out[7] = s->t_peak;
}

long orc_hydrograph(const orc_t *s, double *out, long cap)
{
    long n = s->nhydro < cap ? s->nhydro : cap;
    if (n > 0) memcpy(out, s->hydro, sizeof(double) * 2 * (size_t)n);
    return s->nhydro;
}

void orc_bc_acc(const orc_t *s, double *out)
{
    if (s->nbc > 0) memcpy(out, s->bc_acc, sizeof(double) * (size_t)s->nbc);
}

/* ------------------------------------------------------------------------
 * Coupling: couple_BC_efdc, SW2D-EFDC.cu:728-1067 (volumeFlow_BC == 1).
 * Scans the mask south to north; for every catchment cell (mask != 1) whose
 * 4-neighbour differs it appends a record.  Records are returned in four
 * lists (kind 0 "South", 1 "North", 2 "East", 3 "West" in the reference's
 * naming) with the EFDC (I, J) of the reference and the rate.
 * out arrays are sized >= R*C each; counts[4] receives the list lengths.
 * ------------------------------------------------------------------------ */
int orc_couple_bc(int dimx, int dimy, double dx, double dy, const int *mask, const double *h,
                  const double *vva, const double *uua, int *cells, int *I, int *J,
                  double *q, int *counts, int cap)
{
    /* h_vn receives vva and h_um receives uua at the call sites (4903, 4909) */
    int n[4] = {0, 0, 0, 0};
    for (int ini = 0; ini < dimy; ini++) {
        int posxy = dimy * dimx - (ini + 1) * dimx - 1;
        for (int inj = 0; inj < dimx; inj++) {
            posxy = posxy + 1;
            if (mask[posxy] == 1) continue;
            int interior_row = (ini > 0) && (ini < dimy - 1);
            int chk[4] = {0, 0, 0, 0};  /* S N E W */
            if (interior_row) {
                if (inj > 0 && inj < dimx - 1) { chk[0] = chk[1] = chk[2] = chk[3] = 1; }
                else if (inj == dimx - 1) { chk[0] = chk[1] = chk[2] = 1; }
                else if (inj == 0) { chk[0] = chk[1] = chk[3] = 1; }
            }
            if (inj > 0 && inj < dimx - 1) {
                if (ini == 0) { chk[0] = chk[2] = chk[3] = 1; }
                else if (ini == dimy - 1) { chk[1] = chk[2] = chk[3] = 1; }
            }
            const int nb[4] = {posxy - dimx, posxy + dimx, posxy - 1, posxy + 1};
            const int ii[4] = {inj + 1, inj + 1, inj, inj + 2};
            const int jj[4] = {ini + 2, ini, ini + 1, ini + 1};
            for (int k = 0; k < 4; k++) {
                if (!chk[k] || mask[posxy] == mask[nb[k]]) continue;
                int idx = k * cap + n[k];
                if (n[k] >= cap) return -1;
                cells[idx] = posxy; I[idx] = ii[k]; J[idx] = jj[k];
                q[idx] = orc_bc_rate(k, h[posxy], uua[posxy], vva[posxy], dx, dy);
                n[k]++;
            }
        }
    }
    for (int k = 0; k < 4; k++) counts[k] = n[k];
    return 0;
}

/* lake_catchment_smooth, SW2D-EFDC.cu:2766-2827 (in place on baseo) */
void orc_lake_catchment_smooth(int dimx, int dimy, const int *mask, double *baseo)
{
    for (int ini = 0; ini < dimy; ini++) {
        int posxy = dimy * dimx - (ini + 1) * dimx - 1;
        for (int inj = 0; inj < dimx; inj++) {
            posxy = posxy + 1;
            if (mask[posxy] == 1) continue;
            if (!((ini > 0) && (ini < dimy - 1))) continue;
            if (!(inj > 0 && inj < dimx - 1)) continue;
            if (mask[posxy] != mask[posxy - dimx] && baseo[posxy + dimx] != -9999)
                baseo[posxy] = (baseo[posxy - dimx] + baseo[posxy + dimx]) / 2.0;
            if (mask[posxy] != mask[posxy + dimx] && baseo[posxy - dimx] != -9999)
                baseo[posxy] = (baseo[posxy - dimx] + baseo[posxy + dimx]) / 2.0;
            if (mask[posxy] != mask[posxy - 1] && baseo[posxy + 1] != -9999)
                baseo[posxy] = (baseo[posxy - 1] + baseo[posxy + 1]) / 2.0;
            if (mask[posxy] != mask[posxy + 1] && baseo[posxy - 1] != -9999)
                baseo[posxy] = (baseo[posxy - 1] + baseo[posxy + 1]) / 2.0;
        }
    }
}
