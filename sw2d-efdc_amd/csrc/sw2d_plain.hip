// sw2d_plain.hip -- the reference's unfused FP64 kernel sequence.
//
// One kernel per reference kernel (SW2D-EFDC.cu:97-725), in the order of the
// loop body (4798-4877), over the reference's flat layout: N = rows*cols +
// cols/2 cells (3997) including the half pad row, index c = r*cols + j, and
// all the derived arrays the reference keeps (hm, hn, uu1, vv1, uu, vv, uua,
// vva, umo, vno, ho).  This is the "plain FP64 GPU path": the comparator of
// the fused kernel and a second parity witness.  Differences from the
// reference's launch pattern, all of which leave its arithmetic unchanged:
//   - no host synchronisation between kernels (stream order suffices);
//   - forcing (rain, sources, evaporation) is evaluated on the device from
//     the series instead of a host loop + a dense N-array upload per half step;
//   - the pad row is inactive and a zeroed tail covers the reads past N that
//     the reference makes into unallocated memory;
//   - treat_error is written per face (race-free; the reference's per-cell
//     writes only ever store zeros under sign-exclusive tests).
// Extensions (adaptive step, outlet, storm, Manning map, ledger, lake inflow)
// follow the same definitions as the fused kernel.
#include "sw2d_common.hpp"

namespace sw2d {

struct PlainArgs {
    double *z, *h, *ho, *hm, *hn, *um, *umo, *uu1, *vv1, *uua, *vva, *vn, *vno, *uu, *vv;
    const int *inf;
    const double *ncell;   // per-cell Manning n (ncls_on)
    int N;
    StepCommon s;
};

#define PLAIN_LOOP(c) for (int c = blockIdx.x * blockDim.x + threadIdx.x; c < a.N; c += gridDim.x * blockDim.x)

// flux, SW2D-EFDC.cu:97-353 (+ CFL candidates of the incoming state)
template <bool ADAPT>
__global__ void __launch_bounds__(256) plain_flux(const PlainArgs a)
{
    const StepCommon &s = a.s;
    const StepTime st = step_time(s, iter_of(s));
    const int R = s.R, C = s.C;
    const double th = s.th, gg = s.g, dx = s.dx, dy = s.dy, dt2 = st.dt2;
    const double *h = a.h, *z = a.z, *hm = a.hm, *hn = a.hn, *uu1 = a.uu1, *vv1 = a.vv1;
    const double *umo = a.umo, *vno = a.vno, *uua = a.uua, *vva = a.vva;
    const int *inf = a.inf;
    double cfl = INFINITY;
    PLAIN_LOOP(c) {
        const int inj = c % C, ini = c / C;
        if (ADAPT && ini < R) {
            if (ini > 0 && inf[c] && inf[c - C] && hm[c] > th) {
                double w = dx / (fabs(uu1[c]) + sqrt(gg * hm[c]));
                cfl = (w < cfl) ? w : cfl;
            }
            if (inj > 0 && inf[c] && inf[c - 1] && hn[c] > th) {
                double w = dy / (fabs(vv1[c]) + sqrt(gg * hn[c]));
                cfl = (w < cfl) ? w : cfl;
            }
        }
        double out = 0;
        if (ini > 0 && ini < R && inf[c] != 0 && inf[c - C] != 0 &&
            ((h[c - C] > th) || (h[c] > th))) {
            const double hhe = h[c] + z[c], hhw = h[c - C] + z[c - C];
            const double hhep = h[c] - th, hhwp = h[c - C] - th;
            if (hhe < z[c - C]) {
                out = (h[c - C] > th) ? 0.5440 * h[c - C] * sqrt(gg * h[c - C]) : 0;
            } else if (hhw < z[c]) {
                out = (h[c] > th) ? -0.544 * h[c] * sqrt(gg * h[c]) : 0;
            } else if (hhep * hhwp < 0) {
                if ((h[c] > 0) || (h[c - C] > 0)) {
                    const double hhan = hhep - hhwp;
                    const double sgnm = hhan / fabs(hhan);
                    const double hh3 = fmax((h[c] + z[c]), (h[c - C] + z[c - C])) - fmax(z[c], z[c - C]);
                    out = -sgnm * 0.350 * hh3 * sqrt(2.00 * gg * hh3);
                } else out = 0;
            } else {
                const double u13 = gg * hm[c] * (dt2 / dx) * (h[c] + z[c] - h[c - C] - z[c - C]);
                const double ur = 0.50 * (uu1[c + C] + uu1[c]);
                const double ul = 0.50 * (uu1[c] + uu1[c - C]);
                const double umr = ur * (umo[c + C] + umo[c]) * 0.50 + fabs(ur) * (umo[c] - umo[c + C]) * 0.50;
                const double uml = ul * (umo[c] + umo[c - C]) * 0.50 + fabs(ul) * (umo[c - C] - umo[c]) * 0.50;
                const double u11 = (dt2 / dx) * (umr - uml);
                double u12;
                if (inj == 0) u12 = 0;
                else {
                    const double vu = 0.50 * (vv1[c + 1] + vv1[c - C + 1]);
                    const double vd = 0.50 * (vv1[c] + vv1[c - C]);
                    const double umu = vu * (umo[c + 1] + umo[c]) * 0.50 + fabs(vu) * (umo[c] - umo[c + 1]) * 0.50;
                    const double umd = vd * (umo[c - 1] + umo[c]) * 0.50 + fabs(vd) * (umo[c - 1] - umo[c]) * 0.50;
                    u12 = (dt2 / dy) * (umu - umd);
                }
                const double nf = s.ncls_on ? 0.5 * (a.ncell[c - C] + a.ncell[c]) : s.rn;
                const double sqx = sqrt(uu1[c] * uu1[c] + vva[c] * vva[c]);
                double ram = 0.00;
                if (!(hm[c] <= th)) ram = gg * nf * nf * sqx / pow1333(hm[c]);
                out = ((1.00 - dt2 * ram * 0.50) * umo[c] + (-u11 - u12 - u13)) / (1.00 + dt2 * ram * 0.50);
            }
        }
        a.um[c] = out;
        out = 0;
        if (inj > 0 && inj < C && inf[c] != 0 && inf[c - 1] != 0 &&
            ((h[c - 1] > th) || (h[c] > th))) {
            const double hhn = h[c] + z[c], hhs = h[c - 1] + z[c - 1];
            const double hhnp = h[c] - th, hhsp = h[c - 1] - th;
            if (hhn < z[c - 1]) {
                out = (h[c - 1] > th) ? 0.5440 * h[c - 1] * sqrt(gg * h[c - 1]) : 0;
            } else if (hhs < z[c]) {
                out = (h[c] > th) ? -0.544 * h[c] * sqrt(gg * h[c]) : 0;
            } else if (hhnp * hhsp < 0) {
                if ((h[c] > 0) || (h[c - 1] > 0)) {
                    const double hhan = hhnp - hhsp;
                    const double sgnm = hhan / fabs(hhan);
                    const double hh3 = fmax((h[c] + z[c]), (h[c - 1] + z[c - 1])) - fmax(z[c], z[c - 1]);
                    out = -sgnm * 0.350 * hh3 * sqrt(2.00 * gg * hh3);
                } else out = 0;
            } else {
                const double v13 = gg * hn[c] * (dt2 / dy) * (h[c] + z[c] - h[c - 1] - z[c - 1]);
                double v11;
                if (ini == 0) v11 = 0;
                else if (ini == R) v11 = 0;
                else {
                    const double av = 0.50 * (uu1[c + C] + uu1[c + C - 1]);
                    const double bv = 0.50 * (uu1[c] + uu1[c - 1]);
                    const double vnr = av * (vno[c + C] + vno[c]) * 0.50 + fabs(av) * (vno[c] - vno[c + C]) * 0.50;
                    const double vnl = bv * (vno[c] + vno[c - C]) * 0.50 + fabs(bv) * (vno[c - C] - vno[c]) * 0.50;
                    v11 = (dt2 / dx) * (vnr - vnl);
                }
                const double pv = 0.50 * (vv1[c + 1] + vv1[c]);
                const double qv = 0.50 * (vv1[c] + vv1[c - 1]);
                const double vnu = pv * (vno[c + 1] + vno[c]) * 0.50 + fabs(pv) * (vno[c] - vno[c + 1]) * 0.50;
                const double vnd = qv * (vno[c - 1] + vno[c]) * 0.50 + fabs(qv) * (vno[c - 1] - vno[c]) * 0.50;
                const double v12 = (dt2 / dy) * (vnu - vnd);
                const double nf = s.ncls_on ? 0.5 * (a.ncell[c - 1] + a.ncell[c]) : s.rn;
                const double sqy = sqrt(uua[c] * uua[c] + vv1[c] * vv1[c]);
                double ram = 0.00;
                if (!(hn[c] <= th)) ram = gg * nf * nf * sqy / pow1333(hn[c]);
                out = ((1.00 - dt2 * ram * 0.50) * vno[c] + (-v11 - v12 - v13)) / (1.00 + dt2 * ram * 0.50);
            }
        }
        a.vn[c] = out;
    }
    if (ADAPT) {
        cfl = wave_min(cfl);
        if ((threadIdx.x & 63) == 0 && cfl < INFINITY)
            atomicMin(&s.clk[s.slot_b].cfl_bits, (unsigned long long)__double_as_longlong(cfl));
    }
}

__device__ __forceinline__ double plain_cell_rain(const StepCommon &s, double t, double rr, int c)
{
    if (!s.storm_on) return rr;
    if (!storm_live(s, t)) return 0.0;
    const int r = c / s.C, j = c % s.C;
    const double x = ((double)j + 0.5) * s.dx, y = ((double)r + 0.5) * s.dy;
    const double gx = storm_factor(x - (s.st_x0 + s.st_vx * t), s.st_sigma);
    const double gy = storm_factor(y - (s.st_y0 + s.st_vy * t), s.st_sigma);
    return s.st_peak * (gx * gy);
}

__device__ __forceinline__ double plain_ql(const StepCommon &s, int c, double t)
{
    double ql = 0;
    for (int i = 0; i < s.nst; ++i)
        if (s.src_r[i] * s.C + s.src_c[i] == c) ql = source_rate(s, i, t);
    return ql;
}

// continuity, SW2D-EFDC.cu:602-664.  second = 0: first call (forcing at t);
// second = 1: the call at t + dt, which also writes the ledger partials.
template <int SECOND>
__global__ void __launch_bounds__(256) plain_continuity(const PlainArgs a)
{
    const StepCommon &s = a.s;
    const StepTime st = step_time(s, iter_of(s));
    const double tf = SECOND ? st.tm : st.t;
    const double dt2 = st.dt2, dx = s.dx, dy = s.dy;
    const int C = s.C;
    const double rr0 = rain_rate(s, tf), Ev = evap_rate(s, tf);
    const int oc = s.out_edge ? s.out_r * C + s.out_c : -1;
    double acc_vol = 0, acc_src = 0, acc_clip = 0, acc_q = 0;
    PLAIN_LOOP(c) {
        double P, E;
        if (a.ho[c] > s.EV_WL_min) { P = s.P_lake; E = Ev; }
        else { P = s.P_land; E = 0.000; }
        const double rr = plain_cell_rain(s, tf, rr0, c);
        const double ql = plain_ql(s, c, tf);
        double ums = a.um[c + C], umn = a.um[c], vne = a.vn[c + 1], vnw = a.vn[c];
        double qo = 0;
        if (c == oc) {
            const double hs = a.ho[c];
            qo = (hs > s.th) ? 0.5440 * hs * sqrt(s.g * hs) : 0;
            if (SECOND && ((a.inf[c] == 1) || (a.inf[c] == 2)) && (a.h[c] < s.th)) qo = 0;
            if (s.out_edge == 1) ums = qo;
            else if (s.out_edge == 2) umn = -qo;
            else if (s.out_edge == 3) vnw = -qo;
            else vne = qo;
        }
        const double hr = a.ho[c] - dt2 * ((ums - umn) / dx + (vne - vnw) / dy - ql - rr * P + E);
        double hv = fmax(hr, 0.00);
        if (a.inf[c] == 0) hv = 0.00;
        a.h[c] = hv;
        if (SECOND && a.inf[c]) {
            acc_vol += hv;
            acc_src += dt2 * (ql + rr * P - E);
            if (hr < 0.0) acc_clip += -hr;
            if (c == oc) acc_q += qo * dy;
        }
    }
    if (SECOND) {
        __shared__ double red[4][4];
        const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
        acc_vol = wave_sum(acc_vol); acc_src = wave_sum(acc_src);
        acc_clip = wave_sum(acc_clip); acc_q = wave_sum(acc_q);
        if (lane == 0) { red[wv][0] = acc_vol; red[wv][1] = acc_src; red[wv][2] = acc_clip; red[wv][3] = acc_q; }
        __syncthreads();
        if (threadIdx.x == 0) {
            Partial p;
            p.vol = ((red[0][0] + red[1][0]) + red[2][0]) + red[3][0];
            p.src = ((red[0][1] + red[1][1]) + red[2][1]) + red[3][1];
            p.clip = ((red[0][2] + red[1][2]) + red[2][2]) + red[3][2];
            p.qout = ((red[0][3] + red[1][3]) + red[2][3]) + red[3][3];
            s.part_out[blockIdx.x] = p;
        }
    }
}

// treat_error, SW2D-EFDC.cu:695-725, per face
__global__ void __launch_bounds__(256) plain_treat_error(const PlainArgs a)
{
    const StepCommon &s = a.s;
    const int C = s.C, R = s.R, N = a.N;
    const double th = s.th;
    const double *h = a.h;
    const int *inf = a.inf;
#define PLAIN_DRY(c) (((inf[c] == 1) || (inf[c] == 2)) && (h[c] < th))
    for (int f = blockIdx.x * blockDim.x + threadIdx.x; f < N + C; f += gridDim.x * blockDim.x) {
        if (f >= C && (f - C) / C < R && f - C < N && PLAIN_DRY(f - C) && a.um[f] > 0) a.um[f] = 0;
        if (f < N && PLAIN_DRY(f) && a.um[f] < 0.0) a.um[f] = 0.00;
        if (f >= 1 && f - 1 < N && PLAIN_DRY(f - 1) && a.vn[f] > 0) a.vn[f] = 0;
        if (f < N && PLAIN_DRY(f) && a.vn[f] < 0) a.vn[f] = 0;
    }
#undef PLAIN_DRY
}

// hm_hn, SW2D-EFDC.cu:362-398
__global__ void __launch_bounds__(256) plain_hm_hn(const PlainArgs a)
{
    const int C = a.s.C, R = a.s.R;
    PLAIN_LOOP(c) {
        const int inj = c % C, ini = c / C;
        if (inj > 0) a.hn[c] = 0.50 * (a.h[c] + a.h[c - 1]);
        if (ini > 0) a.hm[c] = 0.50 * (a.h[c] + a.h[c - C]);
        if (ini == 0) a.hm[c] = a.h[c];
        if (ini == R) a.hm[c] = a.h[c - C];
        if (inj == 0) a.hn[c] = a.h[c];
    }
}

// uu1_vv1, SW2D-EFDC.cu:404-452
__global__ void __launch_bounds__(256) plain_uu1_vv1(const PlainArgs a)
{
    const double th = a.s.th;
    PLAIN_LOOP(c) {
        a.uu1[c] = (a.hm[c] >= th) ? a.um[c] / a.hm[c] : 0.00;
        a.vv1[c] = (a.hn[c] >= th) ? a.vn[c] / a.hn[c] : 0.0;
    }
}

// uu_vv, SW2D-EFDC.cu:458-483
__global__ void __launch_bounds__(256) plain_uu_vv(const PlainArgs a)
{
    const double th = a.s.th;
    const int C = a.s.C;
    PLAIN_LOOP(c) {
        if (a.h[c] >= th) {
            a.uu[c] = (a.uu1[c + C] + a.uu1[c]) / 2.00;
            a.vv[c] = (a.vv1[c + 1] + a.vv1[c]) / 2.00;
        } else {
            a.uu[c] = 0.00;
            a.vv[c] = 0.00;
        }
    }
}

// uua_vva, SW2D-EFDC.cu:489-522
__global__ void __launch_bounds__(256) plain_uua_vva(const PlainArgs a)
{
    const int C = a.s.C, R = a.s.R;
    PLAIN_LOOP(c) {
        const int inj = c % C, ini = c / C;
        if (inj > 0) a.uua[c] = 0.250 * (a.uu1[c] + a.uu1[c + C] + a.uu1[c - 1] + a.uu1[c + C - 1]);
        if (ini > 0) a.vva[c] = 0.250 * (a.vv1[c] + a.vv1[c + 1] + a.vv1[c - C] + a.vv1[c - C + 1]);
        if (inj == 0) a.uua[c] = 0.50 * (a.uu1[c] + a.uu1[c + C]);
        if (ini == 0) a.vva[c] = 0.50 * (a.vv1[c] + a.vv1[c + 1]);
        if (ini == R) a.vva[c] = 0.50 * (a.vv1[c] + a.vv1[c + 1]);
    }
}

// forward, SW2D-EFDC.cu:670-693
__global__ void __launch_bounds__(256) plain_forward(const PlainArgs a)
{
    PLAIN_LOOP(c) {
        a.umo[c] = a.um[c];
        a.vno[c] = a.vn[c];
        a.ho[c] = a.h[c];
    }
}

// clock ring, ledger reduction and lake inflow of the incoming state
__global__ void __launch_bounds__(64) plain_bookkeeping(const PlainArgs a)
{
    const StepCommon &s = a.s;
    const StepTime st = step_time(s, iter_of(s));
    const int C = s.C;
    auto vel = [&](int r, int j, double &hh, double &ua, double &va) {
        const int p = r * C + j;
        hh = a.h[p]; ua = a.uua[p]; va = a.vva[p];
    };
    bookkeeping(s, iter_of(s), st, vel, (int)blockIdx.x, threadIdx.x);
}

// init_inf + initiald (SW2D-EFDC.cu:29-91) and the initial CFL minimum
__global__ void __launch_bounds__(256) plain_init(const PlainArgs a, int adaptive)
{
    const int C = a.s.C, R = a.s.R;
    const double th = a.s.th;
    double cfl = INFINITY;
    PLAIN_LOOP(c) {
        const int inj = c % C, ini = c / C;
        a.ho[c] = a.h[c];
        if (ini == 0) a.hm[c] = a.h[c];
        else if (ini == R) a.hm[c] = a.h[c - C];
        else a.hm[c] = 0.50 * (a.h[c] + a.h[c - C]);
        if (inj == 0) a.hn[c] = a.h[c];
        else a.hn[c] = 0.50 * (a.h[c] + a.h[c - 1]);
        if (adaptive && ini < R) {
            // um = vn = 0 initially, so uu1 = vv1 = 0
            if (ini > 0 && a.inf[c] && a.inf[c - C] && a.hm[c] > th) {
                double w = a.s.dx / (0.0 + sqrt(a.s.g * a.hm[c]));
                cfl = (w < cfl) ? w : cfl;
            }
            if (inj > 0 && a.inf[c] && a.inf[c - 1] && a.hn[c] > th) {
                double w = a.s.dy / (0.0 + sqrt(a.s.g * a.hn[c]));
                cfl = (w < cfl) ? w : cfl;
            }
        }
    }
    if (adaptive) {
        cfl = wave_min(cfl);
        if ((threadIdx.x & 63) == 0 && cfl < INFINITY)
            atomicMin(&a.s.clk[0].cfl_bits, (unsigned long long)__double_as_longlong(cfl));
    }
}

static PlainArgs plain_args(sw2d_handle *h, const StepCommon &sc)
{
    PlainArgs a;
    a.z = h->pz; a.h = h->ph; a.ho = h->pho; a.hm = h->phm; a.hn = h->phn;
    a.um = h->pum; a.umo = h->pumo; a.uu1 = h->puu1; a.vv1 = h->pvv1; a.uua = h->puua;
    a.vva = h->pvva; a.vn = h->pvn; a.vno = h->pvno; a.uu = h->puu; a.vv = h->pvv;
    a.inf = h->pinf; a.ncell = h->pncell; a.N = h->N;
    a.s = sc;
    return a;
}

hipError_t launch_plain(sw2d_handle *h, const StepCommon &sc)
{
    const PlainArgs a = plain_args(h, sc);
    const dim3 grid(h->plain_blocks), block(256);
    hipStream_t st = h->stream;
    hipLaunchKernelGGL(plain_bookkeeping, dim3(1 + (h->nbc + 63) / 64), dim3(64), 0, st, a);
    if (h->prm.adaptive) hipLaunchKernelGGL(plain_flux<true>, grid, block, 0, st, a);
    else hipLaunchKernelGGL(plain_flux<false>, grid, block, 0, st, a);
    hipLaunchKernelGGL(plain_continuity<0>, grid, block, 0, st, a);
    hipLaunchKernelGGL(plain_treat_error, grid, block, 0, st, a);
    hipLaunchKernelGGL(plain_continuity<1>, grid, block, 0, st, a);
    hipLaunchKernelGGL(plain_hm_hn, grid, block, 0, st, a);
    hipLaunchKernelGGL(plain_uu1_vv1, grid, block, 0, st, a);
    hipLaunchKernelGGL(plain_uu_vv, grid, block, 0, st, a);
    hipLaunchKernelGGL(plain_uua_vva, grid, block, 0, st, a);
    hipLaunchKernelGGL(plain_forward, grid, block, 0, st, a);
    return hipGetLastError();
}

hipError_t launch_plain_init(sw2d_handle *h)
{
    StepCommon sc{};
    sc.R = h->R; sc.C = h->C; sc.dx = h->grid.cellsize; sc.dy = h->grid.cellsize;
    sc.g = h->prm.g; sc.th = h->prm.th; sc.clk = h->clk;
    const PlainArgs a = plain_args(h, sc);
    hipLaunchKernelGGL(plain_init, dim3(h->plain_blocks), dim3(256), 0, h->stream, a,
                       h->prm.adaptive);
    return hipGetLastError();
}

}  // namespace sw2d
