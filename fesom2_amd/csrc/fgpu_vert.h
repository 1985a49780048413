// Vertical high-order fluxes CDIFF and PPM at interface j (1-based) of one
// column, shared by the per-stage kernels and the fused step. T, W, ar, hn and
// ho point at level/interface 1 of the column (T: the tracer values used for
// the flux, W: interface velocities, ar: interface areas, hn/ho: new/old layer
// thicknesses); nzmin..nzmax are the column's interfaces (nzmax =
// nlevels_nod2D). Every expression keeps the reference's left-to-right
// evaluation order, so the results are bit-identical to a serial run.
#pragma once

#include <hip/hip_runtime.h>

// adv_tra_ver_cdiff (oce_adv_tra_ver.F90:635-695): the bottom interface keeps
// its incoming value
__device__ __forceinline__ double ver_cdiff(int j, int nzmin, int nzmax, const double *T, const double *W,
                                            const double *ar, double f)
{
    if (j == nzmin) return -W[j - 1] * T[j - 1] * ar[j - 1] - f;
    if (j >= nzmin + 1 && j <= nzmax - 1) {
        const double tv = 0.5 * (T[j - 2] + T[j - 1]);
        return -tv * W[j - 1] * ar[j - 1] - f;
    }
    return f;
}

// adv_tra_vert_ppm (oce_adv_tra_ver.F90:438-631): interface value tv(i)
// (Colella and Woodward 1984, eqs. 1.6-1.8, on the new thicknesses), with the
// second-order rows next to the surface and the bottom
__device__ __forceinline__ double ppm_tv(int i, int nzmin, int nzmax, const double *T, const double *hn)
{
    if (i == nzmin) return T[nzmin - 1];
    if (i == nzmax) return T[nzmax - 2];
    if (i == nzmin + 1) return 0.5 * (T[nzmin - 1] + T[nzmin]);
    if (i == nzmax - 1) return 0.5 * (T[nzmax - 3] + T[nzmax - 2]);
    const int nz = i - 1;   // interface between layers nz and nz+1, nzmin+1 <= nz <= nzmax-3
    const double dzjm1 = hn[nz - 2], dzj = hn[nz - 1], dzjp1 = hn[nz], dzjp2 = hn[nz + 1];
    const double tm1 = T[nz - 2], t0 = T[nz - 1], t1 = T[nz], t2 = T[nz + 1];
    double deltaj = dzj / (dzjm1 + dzj + dzjp1) *
                    ((2.0 * dzjm1 + dzj) / (dzjp1 + dzj) * (t1 - t0) + (dzj + 2.0 * dzjp1) / (dzjm1 + dzj) * (t0 - tm1));
    double deltajp1 = dzjp1 / (dzj + dzjp1 + dzjp2) *
                      ((2.0 * dzj + dzjp1) / (dzjp2 + dzjp1) * (t2 - t1) + (dzjp1 + 2.0 * dzjp2) / (dzj + dzjp1) * (t1 - t0));
    if ((t1 - t0) * (t0 - tm1) > 0.0)
        deltaj = fmin(fmin(fabs(deltaj), 2.0 * fabs(t1 - t0)), 2.0 * fabs(t0 - tm1)) * copysign(1.0, deltaj);
    else
        deltaj = 0.0;
    if ((t2 - t1) * (t1 - t0) > 0.0)
        deltajp1 = fmin(fmin(fabs(deltajp1), 2.0 * fabs(t2 - t1)), 2.0 * fabs(t1 - t0)) * copysign(1.0, deltajp1);
    else
        deltajp1 = 0.0;
    return t0 + dzj / (dzj + dzjp1) * (t1 - t0) +
           1.0 / (dzjm1 + dzj + dzjp1 + dzjp2) *
               (2.0 * dzjp1 * dzj / (dzj + dzjp1) *
                    ((dzjm1 + dzj) / (2.0 * dzj + dzjp1) - (dzjp2 + dzjp1) / (2.0 * dzjp1 + dzj)) * (t1 - t0) -
                dzj * (dzjm1 + dzj) / (2.0 * dzj + dzjp1) * deltajp1 +
                dzjp1 * (dzjp1 + dzjp2) / (dzj + 2.0 * dzjp1) * deltaj);
}

// the limited parabola of layer nz: edge values aL (top), aR (bottom), curvature aj
__device__ __forceinline__ void ppm_layer(int nz, int nzmin, int nzmax, const double *T, const double *hn, double &aL,
                                          double &aR, double &aj)
{
    const double t = T[nz - 1];
    aL = ppm_tv(nz, nzmin, nzmax, T, hn);
    aR = ppm_tv(nz + 1, nzmin, nzmax, T, hn);
    if ((aR - t) * (t - aL) <= 0.0) {
        aL = t;
        aR = t;
    }
    if ((aR - aL) * (t - 0.5 * (aL + aR)) > (aR - aL) * (aR - aL) / 6.0) aL = 3.0 * t - 2.0 * aR;
    if ((aR - aL) * (t - 0.5 * (aR + aL)) < -((aR - aL) * (aR - aL) / 6.0)) aR = 3.0 * t - 2.0 * aL;
    aj = 6.0 * (t - 0.5 * (aL + aR));
}

// flux at interface j: the swept part of the layer below (W > 0) or above
// (W < 0) for a Courant number x = min(|W| dt / h_old, 1); surface row
// -tv W area, zero bottom row
__device__ __forceinline__ double ver_ppm(int j, int nzmin, int nzmax, const double *T, const double *W,
                                          const double *ar, const double *hn, const double *ho, double dt, double f)
{
    double tvert = 0.0;
    const double w = W[j - 1];
    if (j == nzmin) {
        tvert = -ppm_tv(nzmin, nzmin, nzmax, T, hn) * w * ar[j - 1];
    } else if (j == nzmax) {
        tvert = 0.0;
    } else if (w > 0.0) {   // layer j, below the interface
        double aL, aR, aj;
        ppm_layer(j, nzmin, nzmax, T, hn, aL, aR, aj);
        const double x = fmin(w * dt / ho[j - 1], 1.0);
        tvert = (-aL - 0.5 * x * (aR - aL + (1.0 - 2.0 / 3.0 * x) * aj));
        tvert = tvert * ar[j - 1] * w;
    } else if (w < 0.0) {   // layer j-1, above the interface
        double aL, aR, aj;
        ppm_layer(j - 1, nzmin, nzmax, T, hn, aL, aR, aj);
        const double x = fmin(-w * dt / ho[j - 2], 1.0);
        tvert = (-aR + 0.5 * x * (aR - aL - (1.0 - 2.0 / 3.0 * x) * aj));
        tvert = tvert * ar[j - 1] * w;
    }
    return tvert - f;
}
