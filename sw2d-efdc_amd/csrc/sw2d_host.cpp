// sw2d_host.cpp -- host-side pieces of the SW2D path that run on the CPU in the
// reference too: the SW2D -> EFDC lake-inflow coupling scan, the lake/catchment
// DEM smoothing and the evaporation series.
#include <math.h>

#include "../../include/sw2d.h"

namespace {

// inflow of one record toward the lake (SW2D-EFDC.cu:764-821, volumeFlow_BC 1)
inline double record_rate(int kind, double h, double uua, double vva, double dx, double dy)
{
    switch (kind) {
    case 0: return (uua < 0) ? -uua * h * dx : 0.00;  // "South": mask differs at p - cols
    case 1: return (uua > 0) ? uua * h * dx : 0.00;   // "North": p + cols
    case 2: return (vva < 0) ? -vva * h * dy : 0.00;  // "East": p - 1
    default: return (vva > 0) ? vva * h * dy : 0.00;  // "West": p + 1
    }
}

}  // namespace

extern "C" {

// couple_BC_efdc, SW2D-EFDC.cu:728-1067.  The scan visits memory rows from the
// last (ini = 0, the south edge) to the first, columns west to east; interior
// cells test all four neighbours, edge columns and edge rows the three that
// exist, corner cells none.  EFDC (I, J) per list: S (inj+1, ini+2),
// N (inj+1, ini), E (inj, ini+1), W (inj+2, ini+1).
int sw2d_couple_bc(int rows, int cols, double dx, double dy, const int *mask, const double *h,
                   const double *vva, const double *uua, int cap, int *cells, int *I, int *J,
                   double *q, int *counts)
{
    if (rows < 1 || cols < 1 || !mask || !h || !vva || !uua || cap < 0 || !cells || !I || !J ||
        !q || !counts)
        return SW2D_E_ARG;
    const int dimx = cols, dimy = rows;
    int n[4] = {0, 0, 0, 0};
    for (int ini = 0; ini < dimy; ini++) {
        const int rowbase = (dimy - 1 - ini) * dimx;
        const bool mid_row = (ini > 0) && (ini < dimy - 1);
        for (int inj = 0; inj < dimx; inj++) {
            const int p = rowbase + inj;
            if (mask[p] == 1) continue;
            const bool mid_col = (inj > 0) && (inj < dimx - 1);
            bool test[4] = {false, false, false, false};
            if (mid_row) {
                test[0] = test[1] = true;
                test[2] = mid_col || inj == dimx - 1;
                test[3] = mid_col || inj == 0;
            } else if (mid_col) {
                if (ini == 0) test[0] = test[2] = test[3] = true;
                else test[1] = test[2] = test[3] = true;
            }
            const int nb[4] = {p - dimx, p + dimx, p - 1, p + 1};
            const int ii[4] = {inj + 1, inj + 1, inj, inj + 2};
            const int jj[4] = {ini + 2, ini, ini + 1, ini + 1};
            for (int k = 0; k < 4; k++) {
                if (!test[k] || mask[p] == mask[nb[k]]) continue;
                if (n[k] >= cap) return SW2D_E_ARG;
                const size_t o = (size_t)k * cap + n[k];
                cells[o] = p;
                I[o] = ii[k];
                J[o] = jj[k];
                q[o] = record_rate(k, h[p], uua[p], vva[p], dx, dy);
                n[k]++;
            }
        }
    }
    for (int k = 0; k < 4; k++) counts[k] = n[k];
    return SW2D_OK;
}

// lake_catchment_smooth, SW2D-EFDC.cu:2766-2827: interior catchment cells next
// to the lake take the mean of the two neighbours across the differing face,
// tested S, N, E, W in that order (later tests overwrite earlier ones).
int sw2d_lake_catchment_smooth(int rows, int cols, const int *mask, double *z)
{
    if (rows < 1 || cols < 1 || !mask || !z) return SW2D_E_ARG;
    const int dimx = cols, dimy = rows;
    for (int ini = 1; ini < dimy - 1; ini++) {
        const int rowbase = (dimy - 1 - ini) * dimx;
        for (int inj = 1; inj < dimx - 1; inj++) {
            const int p = rowbase + inj;
            if (mask[p] == 1) continue;
            if (mask[p] != mask[p - dimx] && z[p + dimx] != -9999) z[p] = (z[p - dimx] + z[p + dimx]) / 2.0;
            if (mask[p] != mask[p + dimx] && z[p - dimx] != -9999) z[p] = (z[p - dimx] + z[p + dimx]) / 2.0;
            if (mask[p] != mask[p - 1] && z[p + 1] != -9999) z[p] = (z[p - 1] + z[p + 1]) / 2.0;
            if (mask[p] != mask[p + 1] && z[p - 1] != -9999) z[p] = (z[p - 1] + z[p + 1]) / 2.0;
        }
    }
    return SW2D_OK;
}

// gpu_evaporation_calc, SW2D-EFDC.cu:556-586 (energy-balance evaporation,
// mm per interval), evaluated once on the host: the series has one value per
// rain interval, far too few to be worth a kernel.
int sw2d_evaporation_series(double albedo, const double *T, const double *Rg, int n, double dtime,
                            double *evapo_mm)
{
    if (n < 0 || (n > 0 && (!T || !Rg || !evapo_mm))) return SW2D_E_ARG;
    for (int i = 0; i < n; i++) {
        const double t = T[i];
        const double Rs = (1 - albedo) * Rg[i];
        const double pw = 999.85308 + 6.32693 * (0.01) * t - 8.523829 * (0.001) * t * t +
                          6.943248 * (0.00001) * t * t * t - 3.821216 * (0.0000001) * t * t * t * t;
        const double tk = t + 273.15;
        const double lv = ((tk / (tk - 33.91)) * (tk / (tk - 33.91))) * 1.91846 * 1000000;
        double e = (Rs / (lv * pw)) * dtime * 1000.0;
        if (e < 0) e = 0;
        evapo_mm[i] = e;
    }
    return SW2D_OK;
}

}  // extern "C"
