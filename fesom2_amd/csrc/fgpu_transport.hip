// Tracer transport kernels for gfx950: AB2/AB3 extrapolation, MFCT/UPW1
// horizontal and QR4C/UPW1 vertical advection, the 3-D FCT limiter,
// flux2dtracer, the T* update and the implicit vertical diffusion.
//
// Layout: the reference's own, level fastest: layer arrays a[n*L + k-1],
// interface arrays a[n*NL + k-1], edge arrays a[e*L + k-1], uv[(el*L+k-1)*2+c].
// Work arrays carry one tracer per slab; gridDim.y runs over tracers, so one
// launch advances every tracer of the batch.
//
// Thread mapping: one thread per (column, level) over the dense index
// col*L + k-1; a wavefront therefore reads contiguous column segments
// (coalesced), and threads of dry levels exit at once. Edge->node sums are
// gathers over the node's edge list in ascending edge order, which is the
// order of the reference's serial scatters: no float atomics, bitwise
// reproducible. Floating-point expressions keep the reference's evaluation
// order and the library is built with -ffp-contract=off.
#include <algorithm>
#include <cstring>

#include "fgpu_internal.h"
#include "fgpu_vert.h"

#define BLK 256

__device__ __forceinline__ double dmx(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmn(double a, double b) { return a < b ? a : b; }

// upwind flux of adv_tra_hor_upw1 (oce_adv_tra_hor.F90:151-157)
__device__ __forceinline__ double upw1_h(double t1, double t2, double vflux, double fin)
{
    return -0.5 * (t1 * (vflux + fabs(vflux)) + t2 * (vflux - fabs(vflux))) - fin;
}

// ------------------------------------------------------------------------------
// K1: init_tracers_AB (oce_tracer_mod.F90:44-123): valuesAB and the valuesold
// rotation, every level of every column.
// ------------------------------------------------------------------------------
__global__ void __launch_bounds__(BLK) k_init_ab(DevMesh m, DevTracers t, int tr0)
{
    const int tr = tr0 + blockIdx.y;
    const size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= t.sN) return;
    const int abo = t.abo[tr];
    double *T = t.T + t.sN * tr;
    double *Told = t.Told + t.sOld * tr;
    double *TAB = t.TAB + t.sN * tr;
    const double v = T[i];
    if (abo == 2) {
        TAB[i] = -(0.5 + FGPU_EPS_AB) * Told[i] + (1.5 + FGPU_EPS_AB) * v;
        Told[i] = v;
    } else {
        double a = 5.0 * Told[2 * i + 1] - 16.0 * Told[2 * i] + 23.0 * v;
        TAB[i] = a / 12.0;
        Told[2 * i + 1] = Told[2 * i];
        Told[2 * i] = v;
    }
}

// ------------------------------------------------------------------------------
// K2: tracer_gradient_elements (oce_tracer_mod.F90:147-188) from values
// ------------------------------------------------------------------------------
__global__ void __launch_bounds__(BLK) k_tr_xy(DevMesh m, DevTracers t, int tr0)
{
    const int tr = tr0 + blockIdx.y;
    const size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= (size_t)m.F2 * m.L) return;
    const int el = (int)(i / m.L);
    const int k = (int)(i - (size_t)el * m.L) + 1;
    if (k < m.ulev[el] || k > m.nlev[el] - 1) return;
    const double *T = t.T + t.sN * tr;
    const int n0 = m.elem_nodes[3 * el], n1 = m.elem_nodes[3 * el + 1], n2 = m.elem_nodes[3 * el + 2];
    const double t0 = T[(size_t)n0 * m.L + k - 1], t1 = T[(size_t)n1 * m.L + k - 1],
                 t2 = T[(size_t)n2 * m.L + k - 1];
    const double *gs = m.gsca + 6 * (size_t)el;
    double *o = t.tr_xy + t.sStg * tr + 2 * i;
    o[0] = gs[0] * t0 + gs[1] * t1 + gs[2] * t2;
    o[1] = gs[3] * t0 + gs[4] * t1 + gs[5] * t2;
}

// ------------------------------------------------------------------------------
// fill_up_dn_grad for one (edge, level) (oce_muscl_adv.F90:356-525), evaluated
// where the MFCT flux needs it instead of materialising edge_up_dn_grad.
// ------------------------------------------------------------------------------
__device__ __forceinline__ void node_mean_grad(const DevMesh &m, const double *trxy, int node, int k,
                                               double &gx, double &gy)
{
    double tvol = 0.0, tx = 0.0, ty = 0.0;
    const int num = m.nie_num[node];
    for (int j = 0; j < num; ++j) {
        const int el = m.nie[(size_t)node * m.kmax + j];
        if (m.nlev[el] - 1 < k || k < m.ulev[el]) continue;
        const double a = m.elem_area[el];
        const double *g = trxy + ((size_t)el * m.L + k - 1) * 2;
        tvol = tvol + a;
        tx = tx + g[0] * a;
        ty = ty + g[1] * a;
    }
    gx = tx / tvol;
    gy = ty / tvol;
}

__device__ __forceinline__ void up_dn_grad(const DevMesh &m, const double *trxy, int e, int n1, int n2, int k,
                                           double g[4])
{
    g[0] = g[1] = g[2] = g[3] = 0.0;
    const int up = m.updn[2 * e], dn = m.updn[2 * e + 1];
    double gx, gy;
    if (up >= 0 && dn >= 0) {
        const int nzmin = max(m.ulevn_max[n1], m.ulevn_max[n2]);
        const int nzmax = min(m.nlevn_min[n1], m.nlevn_min[n2]);
        // writes in the reference's order; a later write wins
        if (k >= m.ulevn[n1] && k <= nzmin - 1) { node_mean_grad(m, trxy, n1, k, gx, gy); g[0] = gx; g[2] = gy; }
        if (k >= m.ulevn[n2] && k <= nzmin - 1) { node_mean_grad(m, trxy, n2, k, gx, gy); g[1] = gx; g[3] = gy; }
        if (k >= nzmin && k <= nzmax - 1) {
            const double *gu = trxy + ((size_t)up * m.L + k - 1) * 2;
            const double *gd = trxy + ((size_t)dn * m.L + k - 1) * 2;
            g[0] = gu[0]; g[1] = gd[0]; g[2] = gu[1]; g[3] = gd[1];
        }
        if (k >= nzmax && k <= m.nlevn[n1] - 1) { node_mean_grad(m, trxy, n1, k, gx, gy); g[0] = gx; g[2] = gy; }
        if (k >= nzmax && k <= m.nlevn[n2] - 1) { node_mean_grad(m, trxy, n2, k, gx, gy); g[1] = gx; g[3] = gy; }
    } else {
        if (k >= m.ulevn[n1] && k <= m.nlevn[n1] - 1) { node_mean_grad(m, trxy, n1, k, gx, gy); g[0] = gx; g[2] = gy; }
        if (k >= m.ulevn[n2] && k <= m.nlevn[n2] - 1) { node_mean_grad(m, trxy, n2, k, gx, gy); g[1] = gx; g[3] = gy; }
    }
}

// debug: materialise edge_up_dn_grad of one tracer (per-function parity)
__global__ void __launch_bounds__(BLK) k_fill_up_dn_grad(DevMesh m, DevTracers t, int tr)
{
    const size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= (size_t)m.E2 * m.L) return;
    const int e = (int)(i / m.L);
    const int k = (int)(i - (size_t)e * m.L) + 1;
    const int n1 = m.edges[2 * e], n2 = m.edges[2 * e + 1];
    if (k > max(m.nlevn[n1], m.nlevn[n2]) - 1) return;
    double g[4];
    up_dn_grad(m, t.tr_xy + t.sStg * tr, e, n1, n2, k, g);
    double *o = t.dbg_grad + 4 * i;
    // levels a node does not reach keep zero in components it would write
    o[0] = g[0]; o[1] = g[1]; o[2] = g[2]; o[3] = g[3];
}

// ------------------------------------------------------------------------------
// K3: horizontal fluxes per (edge, level): LO upwind on T^n
// (adv_tra_hor_upw1, oce_adv_tra_hor.F90:64-257) and the HO flux on T^AB
// (adv_tra_hor_mfct :546-834 or upw1) minus LO = antidiffusive flux.
// The five level segments A-E are applied in the reference's order.
// ------------------------------------------------------------------------------
__device__ __forceinline__ double seg_vflux(int seg, const double *uv1, const double *uv2, double h1, double h2,
                                            double dX1, double dY1, double dX2, double dY2)
{
    if (seg == 0 || seg == 3) return (-uv1[1] * dX1 + uv1[0] * dY1) * h1;
    if (seg == 1 || seg == 4) return (uv2[1] * dX2 - uv2[0] * dY2) * h2;
    return (-uv1[1] * dX1 + uv1[0] * dY1) * h1 + (uv2[1] * dX2 - uv2[0] * dY2) * h2;
}

__global__ void __launch_bounds__(BLK) k_edge_flux(DevMesh m, DevDyn d, DevTracers t, int tr0)
{
    const int tr = tr0 + blockIdx.y;
    const size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= (size_t)m.E2 * m.L) return;
    const int e = (int)(i / m.L);
    const int k = (int)(i - (size_t)e * m.L) + 1;
    const int nl1 = m.einfo[4 * e], nu1 = m.einfo[4 * e + 1], nl2 = m.einfo[4 * e + 2], nu2 = m.einfo[4 * e + 3];
    const int nl12 = min(nl1, nl2), nu12 = max(nu1, nu2);
    bool in[5];
    in[0] = (k >= nu1 && k <= nu12 - 1);
    in[1] = (nu2 > 0 && k >= nu2 && k <= nu12 - 1);
    in[2] = (k >= nu12 && k <= nl12);
    in[3] = (k >= nl12 + 1 && k <= nl1);
    in[4] = (k >= nl12 + 1 && k <= nl2);
    if (!(in[0] || in[1] || in[2] || in[3] || in[4])) return;
    const int n1 = m.edges[2 * e], n2 = m.edges[2 * e + 1];
    const int el1 = m.edge_tri[2 * e], el2 = m.edge_tri[2 * e + 1];
    const double dX1 = m.ecd[4 * e], dY1 = m.ecd[4 * e + 1];
    double dX2 = 0.0, dY2 = 0.0;
    double a = FGPU_R_EARTH * m.elem_cos[el1];
    if (el2 >= 0) {
        dX2 = m.ecd[4 * e + 2];
        dY2 = m.ecd[4 * e + 3];
        a = 0.5 * (a + FGPU_R_EARTH * m.elem_cos[el2]);
    }
    const double *uv1 = d.uv + ((size_t)el1 * m.L + k - 1) * 2;
    const double *uv2 = el2 >= 0 ? d.uv + ((size_t)el2 * m.L + k - 1) * 2 : uv1;
    const double h1 = m.helem[(size_t)el1 * m.L + k - 1];
    const double h2 = el2 >= 0 ? m.helem[(size_t)el2 * m.L + k - 1] : 0.0;
    const int lim = t.lim[tr], hor = t.hor[tr];
    const double *T = t.T + t.sN * tr;
    const double *TAB = t.TAB + t.sN * tr;
    const size_t o1 = (size_t)n1 * m.L + k - 1, o2 = (size_t)n2 * m.L + k - 1;
    double flo = 0.0;
    if (lim == LIM_FCT) {
        const double T1 = T[o1], T2 = T[o2];
#pragma unroll
        for (int s = 0; s < 5; ++s)
            if (in[s]) flo = upw1_h(T1, T2, seg_vflux(s, uv1, uv2, h1, h2, dX1, dY1, dX2, dY2), flo);
        t.lo_h[t.sE * tr + i] = flo;
    }
    const double A1 = TAB[o1], A2 = TAB[o2];
    double f = flo;   // o_init_zero=.false. with FCT, zero otherwise
    if (hor != HOR_UPW1) {   // MFCT, MUSCL
        const double no = t.ph[tr];
        double g[4];
        up_dn_grad(m, t.tr_xy + t.sStg * tr, e, n1, n2, k, g);
        const double edx = m.edge_dxdy[2 * e], edy = m.edge_dxdy[2 * e + 1];
        double Tmean2, Tmean1;
        if (hor == HOR_MUSCL) {   // reconstruction only down to nboundary_lay (oce_adv_tra_hor.F90:356-367)
            const double c1 = k <= m.nblay[n1] ? 1.0 : 0.0, c2 = k <= m.nblay[n2] ? 1.0 : 0.0;
            Tmean2 = A2 - (2.0 * (A2 - A1) + edx * a * g[1] + edy * FGPU_R_EARTH * g[3]) / 6.0 * c2;
            Tmean1 = A1 + (2.0 * (A2 - A1) + edx * a * g[0] + edy * FGPU_R_EARTH * g[2]) / 6.0 * c1;
        } else {
            Tmean2 = A2 - (2.0 * (A2 - A1) + edx * a * g[1] + edy * FGPU_R_EARTH * g[3]) / 6.0;
            Tmean1 = A1 + (2.0 * (A2 - A1) + edx * a * g[0] + edy * FGPU_R_EARTH * g[2]) / 6.0;
        }
#pragma unroll
        for (int s = 0; s < 5; ++s)
            if (in[s]) {
                const double vflux = seg_vflux(s, uv1, uv2, h1, h2, dX1, dY1, dX2, dY2);
                const double cHO = (vflux + fabs(vflux)) * Tmean1 + (vflux - fabs(vflux)) * Tmean2;
                f = -0.5 * (1.0 - no) * cHO - vflux * no * 0.5 * (Tmean1 + Tmean2) - f;
            }
    } else {
#pragma unroll
        for (int s = 0; s < 5; ++s)
            if (in[s]) f = upw1_h(A1, A2, seg_vflux(s, uv1, uv2, h1, h2, dX1, dY1, dX2, dY2), f);
    }
    t.a_h[t.sE * tr + i] = f;
}

// ------------------------------------------------------------------------------
// vertical fluxes at interface j of one column (1-based j), applying the
// reference's special rows in its order. T, W, ar point at level/interface 1.
// ------------------------------------------------------------------------------
// adv_tra_ver_upw1 (oce_adv_tra_ver.F90:244-328)
__device__ __forceinline__ double ver_upw1(int j, int nzmin, int nzmax, const double *T, const double *W,
                                           const double *ar, double f)
{
    if (j == nzmin) f = -W[j - 1] * T[j - 1] * ar[j - 1] - f;
    if (j == nzmax) f = 0.0 - f;
    if (j >= nzmin + 1 && j <= nzmax - 1) {
        const double w = W[j - 1];
        f = -0.5 * (T[j - 1] * (w + fabs(w)) + T[j - 2] * (w - fabs(w))) * ar[j - 1] - f;
    }
    return f;
}

// adv_tra_ver_qr4c (oce_adv_tra_ver.F90:332-434)
__device__ __forceinline__ double ver_qr4c(int j, int nzmin, int nzmax, const double *T, const double *W,
                                           const double *ar, const double *Zn, const double *zb,
                                           double num_ord, double f)
{
    if (j == nzmin) f = -T[j - 1] * W[j - 1] * ar[j - 1] - f;
    if (j == nzmin + 1) f = -0.5 * (T[j - 2] + T[j - 1]) * W[j - 1] * ar[j - 1] - f;
    if (j == nzmax - 1) f = -0.5 * (T[j - 2] + T[j - 1]) * W[j - 1] * ar[j - 1] - f;
    if (j == nzmax) f = 0.0 - f;
    if (j >= nzmin + 2 && j <= nzmax - 2) {
        const int nz = j;
        const double qc = (T[nz - 2] - T[nz - 1]) / (Zn[nz - 2] - Zn[nz - 1]);
        const double qu = (T[nz - 1] - T[nz]) / (Zn[nz - 1] - Zn[nz]);
        const double qd = (T[nz - 3] - T[nz - 2]) / (Zn[nz - 3] - Zn[nz - 2]);
        const double Tmean1 = T[nz - 1] + (2 * qc + qu) * (zb[nz - 1] - Zn[nz - 1]) / 3.0;
        const double Tmean2 = T[nz - 2] + (2 * qc + qd) * (zb[nz - 1] - Zn[nz - 2]) / 3.0;
        const double ww = W[nz - 1];
        const double Tmean = (ww + fabs(ww)) * Tmean1 + (ww - fabs(ww)) * Tmean2;
        f = (-0.5 * (1.0 - num_ord) * Tmean - num_ord * (0.5 * (Tmean1 + Tmean2)) * ww) * ar[nz - 1] - f;
    }
    return f;
}

__device__ __forceinline__ double ver_ho(int ver, int j, int nzmin, int nzmax, const double *TAB, const double *W,
                                         const double *ar, const double *Zn, const double *zb, double pv,
                                         const double *hn, const double *ho, double dt, double f)
{
    if (ver == VER_QR4C) return ver_qr4c(j, nzmin, nzmax, TAB, W, ar, Zn, zb, pv, f);
    if (ver == VER_CDIFF) return ver_cdiff(j, nzmin, nzmax, TAB, W, ar, f);
    if (ver == VER_PPM) return ver_ppm(j, nzmin, nzmax, TAB, W, ar, hn, ho, dt, f);
    return ver_upw1(j, nzmin, nzmax, TAB, W, ar, f);
}

// ------------------------------------------------------------------------------
// K4: per (node, level): LO solution (oce_adv_tra_driver.F90:111-252) and the
// antidiffusive vertical flux at the top interface of the cell (and at the
// bottom interface for the deepest cell).
// ------------------------------------------------------------------------------
__global__ void __launch_bounds__(BLK) k_node_lo(DevMesh m, DevDyn d, DevTracers t, int tr0, double dt)
{
    const int tr = tr0 + blockIdx.y;
    const size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= t.sN) return;
    const int n = (int)(i / m.L);
    const int k = (int)(i - (size_t)n * m.L) + 1;
    const int nzmin = m.ulevn[n], nzmax = m.nlevn[n];
    if (k < nzmin || k > nzmax - 1) return;
    const int L = m.L, NL = m.NL;
    const double *T = t.T + t.sN * tr + (size_t)n * L;
    const double *TAB = t.TAB + t.sN * tr + (size_t)n * L;
    const double *ar = m.area + (size_t)n * NL;
    const double *Zn = m.Z3d + (size_t)n * L;
    const double *zb = m.zbar3d + (size_t)n * NL;
    double *av = t.a_v + t.sNI * tr + (size_t)n * NL;
    const int lim = t.lim[tr], ver = t.ver[tr];
    const double pv = t.pv[tr];
    const double *hn = m.hnode_new + (size_t)n * L, *ho = m.hnode + (size_t)n * L;
    if (lim == LIM_FCT) {
        const double *We = d.w_e + (size_t)n * NL;
        const double *W = d.w + (size_t)n * NL;
        const double *lo_h = t.lo_h + t.sE * tr;
        double acc = 0.0;
        for (int p = m.ne_ptr[n]; p < m.ne_ptr[n + 1]; ++p) {
            const int q = m.ne_idx[p];
            const int e = q >> 1;
            if (k < m.erange[2 * e] || k > m.erange[2 * e + 1]) continue;
            const double fl = lo_h[(size_t)e * L + k - 1];
            acc = (q & 1) ? acc - fl : acc + fl;
        }
        const double lk = ver_upw1(k, nzmin, nzmax, T, We, ar, 0.0);
        const double lk1 = ver_upw1(k + 1, nzmin, nzmax, T, We, ar, 0.0);
        const double av_ = m.areasvol[(size_t)n * NL + k - 1];
        t.fct_lo[i + t.sN * tr] = (T[k - 1] * m.hnode[i] + (acc + (lk - lk1)) * dt / av_) / m.hnode_new[i];
        // with the W split the LO flux handed to the HO scheme is upwind with
        // the full W (oce_adv_tra_driver.F90:323-331); the implicit part enters
        // the LO solution in k_vert_impl
        const double fk = d.wsplit ? ver_upw1(k, nzmin, nzmax, T, W, ar, 0.0) : lk;
        av[k - 1] = ver_ho(ver, k, nzmin, nzmax, TAB, W, ar, Zn, zb, pv, hn, ho, dt, fk);
        if (k == nzmax - 1) {
            const double fk1 = d.wsplit ? ver_upw1(nzmax, nzmin, nzmax, T, W, ar, 0.0) : lk1;
            av[nzmax - 1] = ver_ho(ver, nzmax, nzmin, nzmax, TAB, W, ar, Zn, zb, pv, hn, ho, dt, fk1);
        }
    } else {
        const double *We = d.w_e + (size_t)n * NL;
        av[k - 1] = ver_ho(ver, k, nzmin, nzmax, TAB, We, ar, Zn, zb, pv, hn, ho, dt, 0.0);
        if (k == nzmax - 1) av[nzmax - 1] = ver_ho(ver, nzmax, nzmin, nzmax, TAB, We, ar, Zn, zb, pv, hn, ho, dt, 0.0);
    }
}

// ------------------------------------------------------------------------------
// K4b (use_wsplit only): adv_tra_vert_impl (oce_adv_tra_ver.F90:90-240) on the
// LO solution of the FCT tracers: implicit first-order upwind advection with
// w_i, one tridiagonal solve per column, thread per column. Coefficients are
// formed on the fly in the forward sweep (each depends only on its own level).
// ------------------------------------------------------------------------------
template <int MAXL>
__global__ void __launch_bounds__(64) k_vert_impl(DevMesh m, DevDyn d, DevTracers t, int tr0, double dt)
{
    const int tr = tr0 + blockIdx.y;
    if (t.lim[tr] != LIM_FCT) return;
    const int n = blockIdx.x * 64 + threadIdx.x;
    if (n >= m.N2) return;
    const int L = m.L, NL = m.NL;
    const int nzmin = m.ulevn[n], nzmax = m.nlevn[n];
    double *TF = t.fct_lo + t.sN * tr + (size_t)n * L;
    const double *wi = d.w_i + (size_t)n * NL;
    const double *ar = m.area + (size_t)n * NL;
    const double *avl = m.areasvol + (size_t)n * NL;
    const double *hn = m.hnode_new + (size_t)n * L;
    double cp[MAXL + 1], tp[MAXL + 1];
#define I(kk) ((kk) - 1)
    const double zinv = 1.0 * dt;
    int nz = nzmin;
    double v_adv = zinv * ar[I(nz)] / avl[I(nz)];
    double a = 0.0;
    double b = hn[I(nz)] + wi[I(nz)] * v_adv;
    v_adv = zinv * ar[I(nz + 1)] / avl[I(nz)];
    b = b - dmn(0.0, wi[I(nz + 1)]) * v_adv;
    double c = -dmx(0.0, wi[I(nz + 1)]) * v_adv;
    double Tm, Tk = TF[I(nz)], Tp = TF[I(nz + 1)];
    double r = -(b - hn[I(nz)]) * Tk - c * Tp;
    cp[I(nz)] = c / b;
    tp[I(nz)] = r / b;
    for (nz = nzmin + 1; nz <= nzmax - 1; ++nz) {
        Tm = Tk;
        Tk = Tp;
        v_adv = zinv * ar[I(nz)] / avl[I(nz)];
        a = dmn(0.0, wi[I(nz)]) * v_adv;
        b = hn[I(nz)] + dmx(0.0, wi[I(nz)]) * v_adv;
        if (nz <= nzmax - 2) {
            v_adv = zinv * ar[I(nz + 1)] / avl[I(nz)];
            b = b - dmn(0.0, wi[I(nz + 1)]) * v_adv;
            c = -dmx(0.0, wi[I(nz + 1)]) * v_adv;
            Tp = TF[I(nz + 1)];
            r = -a * Tm - (b - hn[I(nz)]) * Tk - c * Tp;
        } else {
            c = 0.0;
            r = -a * Tm - (b - hn[I(nz)]) * Tk;
        }
        const double mm = b - cp[I(nz - 1)] * a;
        cp[I(nz)] = c / mm;
        tp[I(nz)] = (r - tp[I(nz - 1)] * a) / mm;
    }
    double x = tp[I(nzmax - 1)];
    TF[I(nzmax - 1)] = TF[I(nzmax - 1)] + x;
    for (nz = nzmax - 2; nz >= nzmin; --nz) {
        x = tp[I(nz)] - cp[I(nz)] * x;
        TF[I(nz)] = TF[I(nz)] + x;
    }
#undef I
}

// ------------------------------------------------------------------------------
// K5: FCT a1+a2 (oce_adv_tra_fct.F90:119-178): element bounds with +-1e3 padding
// ------------------------------------------------------------------------------
__global__ void __launch_bounds__(BLK) k_elem_aux(DevMesh m, DevTracers t, int tr0)
{
    const int tr = tr0 + blockIdx.y;
    if (t.lim[tr] != LIM_FCT) return;
    const size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= (size_t)m.F2 * m.L) return;
    const int el = (int)(i / m.L);
    const int k = (int)(i - (size_t)el * m.L) + 1;
    double *o = t.aux + t.sStg * tr + 2 * i;
    if (k <= m.ulev[el] - 1 || k >= m.nlev[el]) {
        o[0] = -1e3;
        o[1] = 1e3;
        return;
    }
    const double *T = t.T + t.sN * tr;
    const double *LO = t.fct_lo + t.sN * tr;
    double mx[3], mn[3];
#pragma unroll
    for (int j = 0; j < 3; ++j) {
        const size_t q = (size_t)m.elem_nodes[3 * el + j] * m.L + k - 1;
        mx[j] = dmx(LO[q], T[q]);
        mn[j] = dmn(LO[q], T[q]);
    }
    o[0] = dmx(dmx(mx[0], mx[1]), mx[2]);
    o[1] = dmn(dmn(mn[0], mn[1]), mn[2]);
}

// ------------------------------------------------------------------------------
// K6: FCT a3 + b1 + b2 per (node, level) (oce_adv_tra_fct.F90:195-405):
// admissible increments, P+/P- from the vertical and the edge gather, R+/R-.
// ------------------------------------------------------------------------------
__device__ __forceinline__ void tvert(const DevMesh &m, const double *aux, int n, int k, double &vmax, double &vmin)
{
    const int *ne = m.nie + (size_t)n * m.kmax;
    const double *a0 = aux + ((size_t)ne[0] * m.L + k - 1) * 2;
    vmax = a0[0];
    vmin = a0[1];
    const int num = m.nie_num[n];
    for (int j = 1; j < num; ++j) {
        const double *a = aux + ((size_t)ne[j] * m.L + k - 1) * 2;
        vmax = dmx(vmax, a[0]);
        vmin = dmn(vmin, a[1]);
    }
}

__global__ void __launch_bounds__(BLK) k_node_limit(DevMesh m, DevTracers t, int tr0, double dt)
{
    const int tr = tr0 + blockIdx.y;
    if (t.lim[tr] != LIM_FCT) return;
    const size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= t.sN) return;
    const int n = (int)(i / m.L);
    const int k = (int)(i - (size_t)n * m.L) + 1;
    const int nzmin = m.ulevn[n], nzmax = m.nlevn[n];
    if (k < nzmin || k > nzmax - 1) return;
    const int L = m.L, NL = m.NL;
    const double *aux = t.aux + t.sStg * tr;
    const double lo = t.fct_lo[t.sN * tr + i];
    double bmax, bmin, vmax, vmin;
    if (k == nzmax - 1 || k == nzmin) {
        tvert(m, aux, n, k, vmax, vmin);
        bmax = vmax - lo;
        bmin = vmin - lo;
    } else {
        double x0, y0, x1, y1, x2, y2;
        tvert(m, aux, n, k - 1, x0, y0);
        tvert(m, aux, n, k, x1, y1);
        tvert(m, aux, n, k + 1, x2, y2);
        bmax = dmx(dmx(x0, x1), x2) - lo;
        bmin = dmn(dmn(y0, y1), y2) - lo;
    }
    const double *av = t.a_v + t.sNI * tr + (size_t)n * NL;
    double p = 0.0, q = 0.0;
    p = p + (dmx(0.0, av[k - 1]) + dmx(0.0, -av[k]));
    q = q + (dmn(0.0, av[k - 1]) + dmn(0.0, -av[k]));
    const double *ah = t.a_h + t.sE * tr;
    for (int pp = m.ne_ptr[n]; pp < m.ne_ptr[n + 1]; ++pp) {
        const int qq = m.ne_idx[pp];
        const int e = qq >> 1;
        if (k < m.erange[2 * e] || k > m.erange[2 * e + 1]) continue;
        const double fl = ah[(size_t)e * L + k - 1];
        if (qq & 1) {
            p = p + dmx(0.0, -fl);
            q = q + dmn(0.0, -fl);
        } else {
            p = p + dmx(0.0, fl);
            q = q + dmn(0.0, fl);
        }
    }
    const double avol = m.areasvol[(size_t)n * NL + k - 1];
    const double hn = m.hnode_new[i];
    double flux = p * dt / avol / hn + 1e-16;
    t.rp[t.sN * tr + i] = dmn(1.0, bmax / flux);
    flux = q * dt / avol / hn - 1e-16;
    t.rm[t.sN * tr + i] = dmn(1.0, bmin / flux);
}

// ------------------------------------------------------------------------------
// K7: FCT b3 horizontal (oce_adv_tra_fct.F90:458-500): limit A_h in place
// ------------------------------------------------------------------------------
__global__ void __launch_bounds__(BLK) k_edge_limit(DevMesh m, DevTracers t, int tr0)
{
    const int tr = tr0 + blockIdx.y;
    if (t.lim[tr] != LIM_FCT) return;
    const size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= (size_t)m.E2 * m.L) return;
    const int e = (int)(i / m.L);
    const int k = (int)(i - (size_t)e * m.L) + 1;
    if (k < m.erange[2 * e] || k > m.erange[2 * e + 1]) return;
    const int n1 = m.edges[2 * e], n2 = m.edges[2 * e + 1];
    const double *rp = t.rp + t.sN * tr, *rm = t.rm + t.sN * tr;
    double *ah = t.a_h + t.sE * tr;
    double ae = 1.0;
    const double flux = ah[i];
    const size_t q1 = (size_t)n1 * m.L + k - 1, q2 = (size_t)n2 * m.L + k - 1;
    if (flux >= 0.0) {
        ae = dmn(ae, rp[q1]);
        ae = dmn(ae, rm[q2]);
    } else {
        ae = dmn(ae, rm[q1]);
        ae = dmn(ae, rp[q2]);
    }
    ah[i] = ae * ah[i];
}

// FCT b3 vertical (oce_adv_tra_fct.F90:425-456) for interface j of a column
__device__ __forceinline__ double av_limited(int j, int nzmin, int nzmax, const double *av, const double *rp,
                                             const double *rm)
{
    const double flux = av[j - 1];
    if (j > nzmax - 1) return flux;   // bottom interface: not limited
    double ae = 1.0;
    if (j == nzmin) {
        if (flux >= 0.0) ae = dmn(ae, rp[j - 1]);
        else ae = dmn(ae, rm[j - 1]);
    } else {
        if (flux >= 0.0) {
            ae = dmn(ae, rm[j - 2]);
            ae = dmn(ae, rp[j - 1]);
        } else {
            ae = dmn(ae, rp[j - 2]);
            ae = dmn(ae, rm[j - 1]);
        }
    }
    return ae * av[j - 1];
}

// ------------------------------------------------------------------------------
// K8: per (node, level): limited vertical flux, oce_tra_adv_flux2dtracer
// (oce_adv_tra_driver.F90:494-646) and, fused, the del_ttf sum and the T*
// update of diff_tracers_ale (oce_ale_tracer.F90:318-322, 742-755).
// mode 0: fused, T <- T*;  mode 1: write del_ttf_advhoriz/advvert (+ limited A_v).
// ------------------------------------------------------------------------------
// explicit diffusion parts of del_ttf at (node n, level k) from the
// precomputed increments (fgpu_diff.hip), in the order of diff_tracers_ale:
// edges of diff_part_hor_redi ascending, diff_ver_part_expl_ale,
// diff_ver_part_redi_expl
__device__ __forceinline__ double add_explicit_diffusion(const DevMesh &m, const DevDyn &d, const DevTracers &t, int tr,
                                                         int n, int k, size_t i, double avol, double dt, double del)
{
    if (d.hdiff) {
        const double *hd = t.hd + t.sE * tr;
        for (int p = m.ne_ptr[n]; p < m.ne_ptr[n + 1]; ++p) {
            const int q = m.ne_idx[p];
            const int e = q >> 1;
            if (k < m.erange[2 * e] || k > m.erange[2 * e + 1]) continue;
            const double c = hd[(size_t)e * m.L + k - 1];
            const double r = (q & 1) ? 0.0 - c : 0.0 + c;
            del = del + r * dt / avol;
        }
    }
    if (!t.ivd[tr]) del = del + t.ve[t.sN * tr + i];
    if (d.redi) del = del + t.vr[t.sN * tr + i];
    return del;
}

__global__ void __launch_bounds__(BLK) k_node_update(DevMesh m, DevDyn d, DevTracers t, int tr0, double dt, int mode)
{
    const int tr = tr0 + blockIdx.y;
    const size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= t.sN) return;
    const int n = (int)(i / m.L);
    const int k = (int)(i - (size_t)n * m.L) + 1;
    const int nzmin = m.ulevn[n], nzmax = m.nlevn[n];
    if (k < nzmin || k > nzmax - 1) return;
    const int L = m.L, NL = m.NL;
    const bool fct = t.lim[tr] == LIM_FCT;
    const double *av = t.a_v + t.sNI * tr + (size_t)n * NL;
    double fk, fk1;
    if (fct) {
        const double *rp = t.rp + t.sN * tr + (size_t)n * L;
        const double *rm = t.rm + t.sN * tr + (size_t)n * L;
        fk = av_limited(k, nzmin, nzmax, av, rp, rm);
        fk1 = av_limited(k + 1, nzmin, nzmax, av, rp, rm);
    } else {
        fk = av[k - 1];
        fk1 = av[k];
    }
    double *T = t.T + t.sN * tr;
    const double Tk = T[i];
    const double avol = m.areasvol[(size_t)n * NL + k - 1];
    double v = 0.0;
    if (fct) v = v - Tk * m.hnode[i] + t.fct_lo[t.sN * tr + i] * m.hnode_new[i];
    v = v + (fk - fk1) * dt / avol;
    double h = 0.0;
    const double *ah = t.a_h + t.sE * tr;
    for (int p = m.ne_ptr[n]; p < m.ne_ptr[n + 1]; ++p) {
        const int q = m.ne_idx[p];
        const int e = q >> 1;
        if (k < m.erange[2 * e] || k > m.erange[2 * e + 1]) continue;
        const double fl = ah[(size_t)e * L + k - 1];
        if (q & 1) h = h - fl * dt / avol;
        else h = h + fl * dt / avol;
    }
    if (mode == 0) {
        double del = 0.0 + h + v;
        del = add_explicit_diffusion(m, d, t, tr, n, k, i, avol, dt, del);
        del = del + Tk * (m.hnode[i] - m.hnode_new[i]);
        T[i] = Tk + del / m.hnode_new[i];
    } else {
        t.dttf_h[i] = h;
        t.dttf_v[i] = v;
        double *dav = t.dbg_av + (size_t)n * NL;
        dav[k - 1] = fk;
        if (k == nzmax - 1) dav[nzmax - 1] = fk1;
    }
}

// T* update from stored del_ttf_adv* (per-tracer diff_tracers_ale entry)
__global__ void __launch_bounds__(BLK) k_tstar(DevMesh m, DevDyn d, DevTracers t, int tr, double dt)
{
    const size_t i = (size_t)blockIdx.x * BLK + threadIdx.x;
    if (i >= t.sN) return;
    const int n = (int)(i / m.L);
    const int k = (int)(i - (size_t)n * m.L) + 1;
    if (k < m.ulevn[n] || k > m.nlevn[n] - 1) return;
    double *T = t.T + t.sN * tr;
    double del = 0.0 + t.dttf_h[i] + t.dttf_v[i];
    del = add_explicit_diffusion(m, d, t, tr, n, k, i, m.areasvol[(size_t)n * m.NL + k - 1], dt, del);
    del = del + T[i] * (m.hnode[i] - m.hnode_new[i]);
    T[i] = T[i] + del / m.hnode_new[i];
}

// ------------------------------------------------------------------------------
// K9: diff_ver_part_impl_ale (oce_ale_tracer.F90:907-1435) per column: Thomas
// solve of the implicit vertical diffusion, bc_surface at the top
// (:1828-2108), then the salinity clamp of solve_tracers_ale (:502-514) for
// the second tracer. Redi off (isredi = 0), no w-split implicit advection.
// ------------------------------------------------------------------------------
template <int MAXL>
__global__ void __launch_bounds__(64) k_vdiff(DevMesh m, DevDyn d, DevTracers t, int tr0, double dt, int do_diff,
                                              int clamp_tracer)
{
    const int tr = tr0 + blockIdx.y;
    const int n = blockIdx.x * 64 + threadIdx.x;
    if (n >= m.N2) return;
    const int L = m.L, NL = m.NL;
    const int nzmin = m.ulevn[n], nzmax = m.nlevn[n];
    double *T = t.T + t.sN * tr + (size_t)n * L;
    if (do_diff == 2 || (do_diff == 1 && t.ivd[tr])) {
        const double *hn = m.hnode_new + (size_t)n * L;
        const double *ar = m.area + (size_t)n * NL;
        const double *avl = m.areasvol + (size_t)n * NL;
        const double *kv = d.Kv + (size_t)n * NL;
        double Z_n[MAXL + 1], cp[MAXL + 1], tp[MAXL + 1], zbn[MAXL + 2];
#define I(kk) ((kk) - 1)
        // zbar_n/Z_n rebuilt bottom-up from hnode_new (oce_ale_tracer.F90:1088-1096)
        double zb = m.zbar_n_bot[n];
        zbn[I(nzmax)] = zb;
        Z_n[I(nzmax - 1)] = zb + hn[I(nzmax - 1)] / 2.0;
        for (int nz = nzmax - 1; nz >= nzmin + 1; --nz) {
            zb = zb + hn[I(nz)];
            zbn[I(nz)] = zb;
            Z_n[I(nz - 1)] = zb + hn[I(nz - 1)] / 2.0;
        }
        // Redi: K33 = Kv + Ki S3^2 interpolated to the interfaces
        // (oce_ale_tracer.F90:1105-1108, 1138-1144, 1181-1184)
        const bool redi = d.redi != 0;
        const double *S = d.slope + (size_t)n * L * 3, *ki = d.Ki + (size_t)n * L;
#define S3(kk) S[((kk) - 1) * 3 + 2]
#define KI(kk) ki[(kk) - 1]
#define TY_BELOW(nz, zi) ((zbn[I((nz) + 1)] - Z_n[I((nz) + 1)]) * (zi) * (S3(nz) * S3(nz)) * KI(nz) + \
                          (Z_n[I(nz)] - zbn[I((nz) + 1)]) * (zi) * (S3((nz) + 1) * S3((nz) + 1)) * KI((nz) + 1))
#define TY_ABOVE(nz, zi) ((zbn[I(nz)] - Z_n[I(nz)]) * (zi) * (S3((nz) - 1) * S3((nz) - 1)) * KI((nz) - 1) + \
                          (Z_n[I((nz) - 1)] - zbn[I(nz)]) * (zi) * (S3(nz) * S3(nz)) * KI(nz))
        // forward sweep: coefficients on the fly, cp/tp kept
        const double zinv = 1.0 * dt;
        int nz = nzmin;
        double zinv2 = 1.0 / (Z_n[I(nz)] - Z_n[I(nz + 1)]);
        double a = 0.0;
        double c = redi ? -(kv[I(nz + 1)] + TY_BELOW(nz, zinv2)) * zinv2 * zinv * ar[I(nz + 1)] / avl[I(nz)]
                        : -(kv[I(nz + 1)]) * zinv2 * zinv * ar[I(nz + 1)] / avl[I(nz)];
        double b = -c + hn[I(nz)];
        // implicit part of the split W (oce_ale_tracer.F90:1118-1129, 1158-1172,
        // 1194-1202); FCT tracers took it in their LO solution
        const bool wimpl = d.wsplit && t.lim[tr] != LIM_FCT;
        const double *wi = d.w_i + (size_t)n * NL;
        if (wimpl) {
            double v_adv = zinv * (ar[I(nz)] / avl[I(nz)]);
            b = b + wi[I(nz)] * v_adv;
            v_adv = zinv * ar[I(nz + 1)] / avl[I(nz)];
            b = b - dmn(0.0, wi[I(nz + 1)]) * v_adv;
            c = c - dmx(0.0, wi[I(nz + 1)]) * v_adv;
        }
        double zinv1 = zinv2;
        double Tm = 0.0, Tk = T[I(nz)], Tp = T[I(nz + 1)];
        double tr_ = -(b - hn[I(nz)]) * Tk - c * Tp;
        const bool hall = bc_every_layer(t.id[tr], d);
        tr_ = tr_ + bc_surface_dev(t.id[tr], dt, d, m.N2, m.L, n, nzmin, Tk);
        cp[I(nz)] = c / b;
        tp[I(nz)] = tr_ / b;
        for (nz = nzmin + 1; nz <= nzmax - 1; ++nz) {
            Tm = Tk;
            Tk = Tp;
            if (nz <= nzmax - 2) {
                zinv2 = 1.0 / (Z_n[I(nz)] - Z_n[I(nz + 1)]);
                if (redi) {
                    a = -(kv[I(nz)] + TY_ABOVE(nz, zinv1)) * zinv1 * zinv * (ar[I(nz)] / avl[I(nz)]);
                    c = -(kv[I(nz + 1)] + TY_BELOW(nz, zinv2)) * zinv2 * zinv * ar[I(nz + 1)] / avl[I(nz)];
                } else {
                    a = -(kv[I(nz)]) * zinv1 * zinv * (ar[I(nz)] / avl[I(nz)]);
                    c = -(kv[I(nz + 1)]) * zinv2 * zinv * ar[I(nz + 1)] / avl[I(nz)];
                }
                b = -a - c + hn[I(nz)];
                zinv1 = zinv2;
                if (wimpl) {
                    double v_adv = zinv * (ar[I(nz)] / avl[I(nz)]);
                    a = a + dmn(0.0, wi[I(nz)]) * v_adv;
                    b = b + dmx(0.0, wi[I(nz)]) * v_adv;
                    v_adv = zinv * ar[I(nz + 1)] / avl[I(nz)];
                    b = b - dmn(0.0, wi[I(nz + 1)]) * v_adv;
                    c = c - dmx(0.0, wi[I(nz + 1)]) * v_adv;
                }
                Tp = T[I(nz + 1)];
                tr_ = -a * Tm - (b - hn[I(nz)]) * Tk - c * Tp;
                if (hall) tr_ = tr_ + bc_surface_dev(t.id[tr], dt, d, m.N2, m.L, n, nz, Tk);
            } else {
                a = redi ? -(kv[I(nz)] + TY_ABOVE(nz, zinv1)) * zinv1 * zinv * (ar[I(nz)] / avl[I(nz)])
                         : -(kv[I(nz)]) * zinv1 * zinv * (ar[I(nz)] / avl[I(nz)]);
                c = 0.0;
                b = -a + hn[I(nz)];
                if (wimpl) {
                    const double v_adv = zinv * (ar[I(nz)] / avl[I(nz)]);
                    a = a + dmn(0.0, wi[I(nz)]) * v_adv;
                    b = b + dmx(0.0, wi[I(nz)]) * v_adv;
                }
                tr_ = -a * Tm - (b - hn[I(nz)]) * Tk;
                if (hall) tr_ = tr_ + bc_surface_dev(t.id[tr], dt, d, m.N2, m.L, n, nz, Tk);
            }
            const double mm = b - cp[I(nz - 1)] * a;
            cp[I(nz)] = c / mm;
            tp[I(nz)] = (tr_ - tp[I(nz - 1)] * a) / mm;
        }
#undef TY_BELOW
#undef TY_ABOVE
#undef S3
#undef KI
        // back substitution and update
        double x = tp[I(nzmax - 1)];
        T[I(nzmax - 1)] = T[I(nzmax - 1)] + x;
        for (nz = nzmax - 2; nz >= nzmin; --nz) {
            x = tp[I(nz)] - cp[I(nz)] * x;
            T[I(nz)] = T[I(nz)] + x;
        }
#undef I
    }
    if (tr == clamp_tracer) {
        for (int nz = nzmin; nz <= nzmax - 1; ++nz) {
            double s = T[nz - 1];
            if (s > 45.0) s = 45.0;
            if (s < 3.0) s = 3.0;
            T[nz - 1] = s;
        }
    }
}

// ------------------------------------------------------------------------------
// deterministic content sum: sum_n sum_k areasvol*hnode_new*T
// ------------------------------------------------------------------------------
__global__ void __launch_bounds__(BLK) k_content_partial(DevMesh m, const double *T, double *part)
{
    __shared__ double sh[BLK];
    double s = 0.0;
    const size_t total = (size_t)m.N2 * m.L;
    for (size_t i = (size_t)blockIdx.x * BLK + threadIdx.x; i < total; i += (size_t)gridDim.x * BLK) {
        const int n = (int)(i / m.L);
        const int k = (int)(i - (size_t)n * m.L) + 1;
        if (k < m.ulevn[n] || k > m.nlevn[n] - 1) continue;
        s += m.areasvol[(size_t)n * m.NL + k - 1] * m.hnode_new[i] * T[i];
    }
    sh[threadIdx.x] = s;
    __syncthreads();
    for (int w = BLK / 2; w > 0; w >>= 1) {
        if (threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) part[blockIdx.x] = sh[0];
}

__global__ void k_content_final(const double *part, int np, double *out)
{
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        double s = 0.0;
        for (int i = 0; i < np; ++i) s += part[i];
        *out = s;
    }
}

// ==============================================================================
// host side
// ==============================================================================
static inline unsigned nblk(size_t n, int b = BLK) { return (unsigned)((n + b - 1) / b); }

static int check_ready(fgpu_context *c)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    if (!c->vert_ready) return fgpu_fail(FGPU_ERR_STATE, "vertical geometry not set (fgpu_set_vertical)");
    if (!c->vel_ready) return fgpu_fail(FGPU_ERR_STATE, "velocities not set (fgpu_set_velocity)");
    if (!c->tracers_ready) return fgpu_fail(FGPU_ERR_STATE, "tracers not initialised");
    FGPU_HIP(hipSetDevice(c->device));
    return FGPU_OK;
}

static int check_tracer(fgpu_context *c, int tr_num)
{
    if (tr_num < 1 || tr_num > c->t.ntr) return fgpu_fail(FGPU_ERR_ARG, "tr_num out of range");
    if (!c->h_configured[tr_num - 1]) return fgpu_fail(FGPU_ERR_STATE, "tracer not configured");
    return FGPU_OK;
}

#define LAUNCH_CHECK(name)                                                                       \
    do {                                                                                         \
        hipError_t _e = hipGetLastError();                                                       \
        if (_e != hipSuccess) return fgpu_fail(FGPU_ERR_HIP, std::string(name) + ": " + hipGetErrorString(_e)); \
    } while (0)

static int launch_vdiff(fgpu_context *c, int tr0, int ntr, double dt, int do_diff, int clamp_tracer)
{
    dim3 g(nblk(c->N2, 64), ntr);
    const int L = c->nl - 1;
    ProfScope ps(c, "k_vdiff");
    if (L <= 32) hipLaunchKernelGGL(k_vdiff<32>, g, dim3(64), 0, c->stream, c->m, c->d, c->t, tr0, dt, do_diff, clamp_tracer);
    else if (L <= 48) hipLaunchKernelGGL(k_vdiff<48>, g, dim3(64), 0, c->stream, c->m, c->d, c->t, tr0, dt, do_diff, clamp_tracer);
    else if (L <= 72) hipLaunchKernelGGL(k_vdiff<72>, g, dim3(64), 0, c->stream, c->m, c->d, c->t, tr0, dt, do_diff, clamp_tracer);
    else hipLaunchKernelGGL(k_vdiff<FGPU_MAX_NL>, g, dim3(64), 0, c->stream, c->m, c->d, c->t, tr0, dt, do_diff, clamp_tracer);
    LAUNCH_CHECK("k_vdiff");
    return FGPU_OK;
}

static int launch_init_ab(fgpu_context *c, int tr0, int ntr)
{
    const DevMesh &m = c->m;
    {
        ProfScope ps(c, "k_init_ab");
        hipLaunchKernelGGL(k_init_ab, dim3(nblk(c->t.sN), ntr), dim3(BLK), 0, c->stream, m, c->t, tr0);
        LAUNCH_CHECK("k_init_ab");
    }
    {
        ProfScope ps(c, "k_tr_xy");
        hipLaunchKernelGGL(k_tr_xy, dim3(nblk((size_t)m.F2 * m.L), ntr), dim3(BLK), 0, c->stream, m, c->t, tr0);
        LAUNCH_CHECK("k_tr_xy");
    }
    return FGPU_OK;
}

static int launch_adv(fgpu_context *c, int tr0, int ntr, double dt, int mode)
{
    const DevMesh &m = c->m;
    const size_t nE = (size_t)m.E2 * m.L, nF = (size_t)m.F2 * m.L, nN = c->t.sN;
    {
        ProfScope ps(c, "k_edge_flux");
        hipLaunchKernelGGL(k_edge_flux, dim3(nblk(nE), ntr), dim3(BLK), 0, c->stream, m, c->d, c->t, tr0);
        LAUNCH_CHECK("k_edge_flux");
    }
    {
        ProfScope ps(c, "k_node_lo");
        hipLaunchKernelGGL(k_node_lo, dim3(nblk(nN), ntr), dim3(BLK), 0, c->stream, m, c->d, c->t, tr0, dt);
        LAUNCH_CHECK("k_node_lo");
    }
    if (c->d.wsplit) {
        ProfScope ps(c, "k_vert_impl");
        const dim3 g(nblk(c->N2, 64), ntr);
        const int L = c->nl - 1;
        if (L <= 32) hipLaunchKernelGGL(k_vert_impl<32>, g, dim3(64), 0, c->stream, m, c->d, c->t, tr0, dt);
        else if (L <= 48) hipLaunchKernelGGL(k_vert_impl<48>, g, dim3(64), 0, c->stream, m, c->d, c->t, tr0, dt);
        else if (L <= 72) hipLaunchKernelGGL(k_vert_impl<72>, g, dim3(64), 0, c->stream, m, c->d, c->t, tr0, dt);
        else hipLaunchKernelGGL(k_vert_impl<FGPU_MAX_NL>, g, dim3(64), 0, c->stream, m, c->d, c->t, tr0, dt);
        LAUNCH_CHECK("k_vert_impl");
    }
    {
        ProfScope ps(c, "k_elem_aux");
        hipLaunchKernelGGL(k_elem_aux, dim3(nblk(nF), ntr), dim3(BLK), 0, c->stream, m, c->t, tr0);
        LAUNCH_CHECK("k_elem_aux");
    }
    {
        ProfScope ps(c, "k_node_limit");
        hipLaunchKernelGGL(k_node_limit, dim3(nblk(nN), ntr), dim3(BLK), 0, c->stream, m, c->t, tr0, dt);
        LAUNCH_CHECK("k_node_limit");
    }
    {
        ProfScope ps(c, "k_edge_limit");
        hipLaunchKernelGGL(k_edge_limit, dim3(nblk(nE), ntr), dim3(BLK), 0, c->stream, m, c->t, tr0);
        LAUNCH_CHECK("k_edge_limit");
    }
    if (mode == 0) {   // the explicit diffusion terms read T^n, which k_node_update overwrites
        int rc = fgpu_launch_diff(c, tr0, ntr, dt);
        if (rc) return rc;
    }
    {
        ProfScope ps(c, "k_node_update");
        hipLaunchKernelGGL(k_node_update, dim3(nblk(nN), ntr), dim3(BLK), 0, c->stream, m, c->d, c->t, tr0, dt, mode);
        LAUNCH_CHECK("k_node_update");
    }
    return FGPU_OK;
}

extern "C" int fgpu_init_tracers_AB(fgpu_context *c, int tr_num)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_tracer(c, tr_num))) return rc;
    if ((rc = fgpu_ensure_staged(c, 1))) return rc;
    const int tr = tr_num - 1;
    if ((rc = launch_init_ab(c, tr, 1))) return rc;
    hipLaunchKernelGGL(k_fill_up_dn_grad, dim3(nblk((size_t)c->m.E2 * c->m.L)), dim3(BLK), 0, c->stream, c->m, c->t, tr);
    LAUNCH_CHECK("k_fill_up_dn_grad");
    FGPU_HIP(hipStreamSynchronize(c->stream));
    return FGPU_OK;
}

// valuesAB of a tracer computed by the caller's init_tracers_AB; the element
// gradients of values (tr_xy, oce_tracer_mod.F90:126-131), which the same
// routine leaves for fill_up_dn_grad, are recomputed from the resident values
extern "C" int fgpu_tracer_set_valuesAB(fgpu_context *c, int tr_num, const double *valuesAB)
{
    if (!c || !valuesAB) return fgpu_fail(FGPU_ERR_ARG, "null argument");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    if (!c->tracers_ready) return fgpu_fail(FGPU_ERR_STATE, "tracers not initialised");
    int rc;
    if ((rc = check_tracer(c, tr_num))) return rc;
    FGPU_HIP(hipSetDevice(c->device));
    if ((rc = fgpu_ensure_staged(c, 1))) return rc;
    const int tr = tr_num - 1;
    FGPU_HIP(hipMemcpyAsync(c->t.TAB + c->t.sN * tr, valuesAB, c->t.sN * sizeof(double), hipMemcpyHostToDevice,
                            c->stream));
    hipLaunchKernelGGL(k_tr_xy, dim3(nblk((size_t)c->m.F2 * c->m.L), 1), dim3(BLK), 0, c->stream, c->m, c->t, tr);
    LAUNCH_CHECK("k_tr_xy");
    FGPU_HIP(hipStreamSynchronize(c->stream));
    return FGPU_OK;
}

extern "C" int fgpu_do_oce_adv_tra(fgpu_context *c, double dt, int tr_num)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_tracer(c, tr_num))) return rc;
    if ((rc = fgpu_ensure_staged(c, 1))) return rc;
    const int tr = tr_num - 1;
    FGPU_HIP(hipMemsetAsync(c->t.dttf_h, 0, c->t.sN * sizeof(double), c->stream));
    FGPU_HIP(hipMemsetAsync(c->t.dttf_v, 0, c->t.sN * sizeof(double), c->stream));
    FGPU_HIP(hipMemsetAsync(c->t.dbg_av, 0, c->t.sNI * sizeof(double), c->stream));
    if ((rc = launch_adv(c, tr, 1, dt, 1))) return rc;
    FGPU_HIP(hipStreamSynchronize(c->stream));
    return FGPU_OK;
}

extern "C" int fgpu_diff_tracers_ale(fgpu_context *c, double dt, int tr_num)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_tracer(c, tr_num))) return rc;
    if ((rc = fgpu_ensure_staged(c, 1))) return rc;
    const int tr = tr_num - 1;
    // implicit or explicit vertical diffusion: Kv and the surface forcing either way
    if (!c->kv_ready) return fgpu_fail(FGPU_ERR_STATE, "Kv not set (fgpu_set_kv)");
    if (!c->forcing_ready) return fgpu_fail(FGPU_ERR_STATE, "surface forcing not set");
    if ((rc = fgpu_check_bc_inputs(c, tr, 1))) return rc;
    if ((rc = fgpu_ensure_diff(c))) return rc;
    if ((rc = fgpu_launch_diff(c, tr, 1, dt))) return rc;
    hipLaunchKernelGGL(k_tstar, dim3(nblk(c->t.sN)), dim3(BLK), 0, c->stream, c->m, c->d, c->t, tr, dt);
    LAUNCH_CHECK("k_tstar");
    if (c->h_ivd[tr])
        if ((rc = launch_vdiff(c, tr, 1, dt, 1, -1))) return rc;
    FGPU_HIP(hipStreamSynchronize(c->stream));
    return FGPU_OK;
}

extern "C" int fgpu_diff_ver_part_impl_ale(fgpu_context *c, double dt, int tr_num)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_tracer(c, tr_num))) return rc;
    if ((rc = fgpu_ensure_staged(c, 1))) return rc;
    if (!c->kv_ready) return fgpu_fail(FGPU_ERR_STATE, "Kv not set (fgpu_set_kv)");
    if (!c->forcing_ready) return fgpu_fail(FGPU_ERR_STATE, "surface forcing not set");
    const int tr = tr_num - 1;
    if ((rc = fgpu_check_bc_inputs(c, tr, 1))) return rc;
    // diff_ver_part_impl_ale runs regardless of i_vert_diff when called directly
    if ((rc = launch_vdiff(c, tr, 1, dt, 2, -1))) return rc;
    FGPU_HIP(hipStreamSynchronize(c->stream));
    return FGPU_OK;
}

// one batched step of every tracer, enqueued on the context stream
static int enqueue_step(fgpu_context *c, double dt)
{
    if (c->fused) return fgpu_enqueue_fused_step(c, dt);
    const int nt = c->t.ntr;
    int rc;
    if ((rc = launch_init_ab(c, 0, nt))) return rc;
    if ((rc = launch_adv(c, 0, nt, dt, 0))) return rc;
    const int clamp = nt >= 2 ? 1 : -1;
    if ((rc = launch_vdiff(c, 0, nt, dt, 1, clamp))) return rc;
    return FGPU_OK;
}

static int check_step(fgpu_context *c)
{
    for (int i = 0; i < c->t.ntr; ++i)
        if (!c->h_configured[i]) return fgpu_fail(FGPU_ERR_STATE, "tracer " + std::to_string(i + 1) + " not configured");
    if (!c->kv_ready) return fgpu_fail(FGPU_ERR_STATE, "Kv not set (fgpu_set_kv)");
    if (!c->forcing_ready) return fgpu_fail(FGPU_ERR_STATE, "surface forcing not set");
    return fgpu_check_bc_inputs(c, 0, c->t.ntr);
}

static int run_steps(fgpu_context *c, double dt, int nsteps)
{
    int rc;
    if (!c->fused && (rc = fgpu_ensure_staged(c, c->t.ntr))) return rc;
    if ((rc = fgpu_ensure_diff(c))) return rc;   // before any capture
    if (c->prof) {
        for (int s = 0; s < nsteps; ++s)
            if ((rc = enqueue_step(c, dt))) return rc;
        return fgpu_prof_flush(c);
    }
    if (!c->graph_valid || c->graph_dt != dt) {
        if (c->graph) { (void)hipGraphExecDestroy(c->graph); c->graph = nullptr; }
        hipGraph_t g;
        FGPU_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = enqueue_step(c, dt);
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) return rc;
        if (e != hipSuccess) return fgpu_fail(FGPU_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
        e = hipGraphInstantiate(&c->graph, g, nullptr, nullptr, 0);
        (void)hipGraphDestroy(g);
        if (e != hipSuccess) return fgpu_fail(FGPU_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
        c->graph_valid = true;
        c->graph_dt = dt;
    }
    for (int s = 0; s < nsteps; ++s) FGPU_HIP(hipGraphLaunch(c->graph, c->stream));
    return FGPU_OK;
}

extern "C" int fgpu_solve_tracers_ale(fgpu_context *c, double dt, int nsteps)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_step(c))) return rc;
    if (nsteps < 1) return fgpu_fail(FGPU_ERR_ARG, "nsteps < 1");
    if ((rc = run_steps(c, dt, nsteps))) return rc;
    FGPU_HIP(hipStreamSynchronize(c->stream));
    return FGPU_OK;
}

extern "C" int fgpu_solve_tracers_ale_timed(fgpu_context *c, double dt, int nsteps, double *ms)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_step(c))) return rc;
    if (nsteps < 1 || !ms) return fgpu_fail(FGPU_ERR_ARG, "bad arguments");
    hipEvent_t e0, e1;
    FGPU_HIP(hipEventCreate(&e0));
    FGPU_HIP(hipEventCreate(&e1));
    FGPU_HIP(hipEventRecord(e0, c->stream));
    rc = run_steps(c, dt, nsteps);
    FGPU_HIP(hipEventRecord(e1, c->stream));
    FGPU_HIP(hipEventSynchronize(e1));
    float f = 0.f;
    FGPU_HIP(hipEventElapsedTime(&f, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    *ms = f;
    return FGPU_OK;
}

// Autotuner of the fused step's launch shape: each candidate number of tracers
// per block (1, 2, 4, ... and all), crossed with 2, 4 or 8 tracers per Thomas
// chunk and 1, 4 or 8 columns per block in the last kernel, is timed over
// `nsteps` replayed steps and the fastest is kept. Every candidate computes the same arithmetic per column, so
// results do not depend on the choice. The tracer state is saved before and
// restored after the trials.
extern "C" int fgpu_autotune_step(fgpu_context *c, double dt, int nsteps, int *best_tpb, double *best_ms)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if ((rc = check_step(c))) return rc;
    if (nsteps < 1) return fgpu_fail(FGPU_ERR_ARG, "nsteps < 1");
    if (!c->fused) return fgpu_fail(FGPU_ERR_STATE, "the autotuner tunes the fused step (FGPU_STEP=staged is set)");
    const int nt = c->t.ntr;
    const size_t nT = c->t.sN * (size_t)nt, nO = c->t.sOld * (size_t)nt;
    double *save = nullptr;
    FGPU_HIP(hipStreamSynchronize(c->stream));
    FGPU_HIP(hipMalloc((void **)&save, sizeof(double) * (nT + nO)));
    auto restore = [&]() -> int {
        FGPU_HIP(hipMemcpyAsync(c->t.T, save, sizeof(double) * nT, hipMemcpyDeviceToDevice, c->stream));
        FGPU_HIP(hipMemcpyAsync(c->t.Told, save + nT, sizeof(double) * nO, hipMemcpyDeviceToDevice, c->stream));
        FGPU_HIP(hipStreamSynchronize(c->stream));
        return FGPU_OK;
    };
    FGPU_HIP(hipMemcpyAsync(save, c->t.T, sizeof(double) * nT, hipMemcpyDeviceToDevice, c->stream));
    FGPU_HIP(hipMemcpyAsync(save + nT, c->t.Told, sizeof(double) * nO, hipMemcpyDeviceToDevice, c->stream));
    struct Shape { int tpb, tch, cpb, kd; };
    std::vector<Shape> cand;
    const bool narrow = c->m.NL <= 64;   // 64-lane columns: several per block are possible
    // the variants of the last kernel; with extra terms only the lock-step one exists
    std::vector<int> kds = {KD_LOCKSTEP};
    if (!fgpu_kd_extras(c)) kds = {KD_LOCKSTEP, KD_SWEEPWAVE, KD_WAVELOCAL};
    // columns over 64 levels: several per block only for the one-wave-per-column
    // K_D, up to 4 with the short LDS rows, else 2
    const int deep_cpb = c->m.NL <= KDP_SHORT_NL ? 4 : 2;
    for (int kd : kds)
        for (int cpb : {1, 2, 3, 4, 8}) {
            if (narrow ? (cpb == 2 || cpb == 3) : cpb > (kd != KD_LOCKSTEP ? deep_cpb : 1)) continue;
            for (int tch : {2, 4, 5, 8}) {
                // chunks of 5 only with one column per block (elsewhere they run as 2)
                if ((tch == 4 && nt < 3) || (tch == 5 && (nt < 5 || cpb > 1)) || (tch == 8 && nt < 5) ||
                    (tch >= 4 && cpb == 8) || (tch == 8 && cpb > 1) ||
                    (!narrow && tch >= 4 && cpb >= 3))
                    continue;
                for (int v = 1; v < nt; v *= 2) cand.push_back({v, tch, cpb, kd});
                cand.push_back({nt, tch, cpb, kd});
            }
        }
    const bool was_prof = c->prof;
    c->prof = false;
    int best = c->step_tpb, best_tch = c->step_tch, best_cpb = c->step_cpb, best_kd = c->step_kdp;
    float best_t = -1.f;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    rc = FGPU_OK;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess)
        rc = fgpu_fail(FGPU_ERR_HIP, "hipEventCreate failed");
    for (size_t i = 0; i < cand.size() && !rc; ++i) {
        c->step_tpb = cand[i].tpb;
        c->step_tch = cand[i].tch;
        c->step_cpb = cand[i].cpb;
        c->step_kdp = cand[i].kd;
        c->graph_valid = false;
        if ((rc = run_steps(c, dt, 1))) break;   // capture and warm up
        if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fgpu_fail(FGPU_ERR_HIP, "hipEventRecord"); break; }
        if ((rc = run_steps(c, dt, nsteps))) break;
        float ms = 0.f;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fgpu_fail(FGPU_ERR_HIP, "timing the autotuner trial failed");
            break;
        }
        if (best_t < 0.f || ms < best_t) {
            best_t = ms;
            best = cand[i].tpb;
            best_tch = cand[i].tch;
            best_cpb = cand[i].cpb;
            best_kd = cand[i].kd;
        }
    }
    // then the bound kernel K_C with the chosen shape: per column over its
    // levels (0) or over (tracer, level) items with 1, 2 or 4 tracers per block
    int best_kcf = c->step_kcf;
    if (!rc && best_t >= 0.f) {
        c->step_tpb = best;
        c->step_tch = best_tch;
        c->step_cpb = best_cpb;
        c->step_kdp = best_kd;
        for (int kcf : {0, 1, 2, 4}) {
            if (kcf == best_kcf) continue;   // timed above
            c->step_kcf = kcf;
            c->graph_valid = false;
            if ((rc = run_steps(c, dt, 1))) break;
            if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fgpu_fail(FGPU_ERR_HIP, "hipEventRecord"); break; }
            if ((rc = run_steps(c, dt, nsteps))) break;
            float ms = 0.f;
            if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
                rc = fgpu_fail(FGPU_ERR_HIP, "timing the autotuner trial failed");
                break;
            }
            if (ms < best_t) {
                best_t = ms;
                best_kcf = kcf;
            }
        }
    }
    c->step_kcf = best_kcf;
    // then K_A and K_B with two levels per lane, where the rows allow it: each
    // flag flipped once, kept if faster
    for (int *flag : {&c->step_ka2, &c->step_kb2}) {
        if (rc || best_t < 0.f || c->m.L % 2 != 0) break;
        const int tried = *flag;
        *flag = !tried;
        c->graph_valid = false;
        float ms = 0.f;
        if ((rc = run_steps(c, dt, 1))) break;
        if (hipEventRecord(e0, c->stream) != hipSuccess) { rc = fgpu_fail(FGPU_ERR_HIP, "hipEventRecord"); break; }
        if ((rc = run_steps(c, dt, nsteps))) break;
        if (hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
            hipEventElapsedTime(&ms, e0, e1) != hipSuccess) {
            rc = fgpu_fail(FGPU_ERR_HIP, "timing the autotuner trial failed");
            break;
        }
        if (ms < best_t) best_t = ms;
        else *flag = tried;
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    const int rr = restore();
    (void)hipFree(save);
    c->prof = was_prof;
    c->step_tpb = best;
    c->step_tch = best_tch;
    c->step_cpb = best_cpb;
    c->step_kdp = best_kd;
    c->graph_valid = false;
    if (rc) return rc;
    if (rr) return rr;
    if (best_tpb) *best_tpb = best;
    if (best_ms) *best_ms = best_t / nsteps;
    return FGPU_OK;
}

extern "C" int fgpu_set_step_shape(fgpu_context *c, int tracers_per_block, int thomas_chunk, int columns_per_block,
                                   int kd_variant, int kc_tracers, int ka_levels_per_lane, int kb_levels_per_lane)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (tracers_per_block < 1 || kd_variant < KD_LOCKSTEP || kd_variant > KD_WAVELOCAL || kc_tracers < 0 ||
        ka_levels_per_lane < 1 || ka_levels_per_lane > 2 || kb_levels_per_lane < 1 || kb_levels_per_lane > 2)
        return fgpu_fail(FGPU_ERR_ARG, "tracers_per_block >= 1, kd_variant in 0..2, kc_tracers >= 0 and "
                                       "ka_levels_per_lane, kb_levels_per_lane 1 or 2");
    c->step_ka2 = ka_levels_per_lane == 2;
    c->step_kb2 = kb_levels_per_lane == 2;
    c->step_tpb = tracers_per_block;
    c->step_tch = thomas_chunk;
    c->step_cpb = columns_per_block;
    c->step_kdp = kd_variant;
    c->step_kcf = kc_tracers >= 4 ? 4 : (kc_tracers >= 2 ? 2 : kc_tracers);
    c->graph_valid = false;
    return FGPU_OK;
}

extern "C" int fgpu_step_variant(fgpu_context *c, int *kd_variant, int *kc_tracers, int *ka_levels_per_lane,
                                 int *kb_levels_per_lane)
{
    if (!c || !kd_variant || !kc_tracers || !ka_levels_per_lane || !kb_levels_per_lane)
        return fgpu_fail(FGPU_ERR_ARG, "null argument");
    *ka_levels_per_lane = (c->step_ka2 && c->m.L % 2 == 0) ? 2 : 1;
    bool xv = false;
    for (int v : c->h_ver) xv = xv || v == VER_CDIFF || v == VER_PPM;
    *kb_levels_per_lane = (c->step_kb2 && c->m.L % 2 == 0 && !c->d.wsplit && !xv) ? 2 : 1;
    *kd_variant = fgpu_kd_extras(c) ? KD_LOCKSTEP : c->step_kdp;
    *kc_tracers = c->step_kcf >= 4 ? 4 : (c->step_kcf >= 2 ? 2 : (c->step_kcf >= 1 ? 1 : 0));
    return FGPU_OK;
}

extern "C" int fgpu_step_shape(fgpu_context *c, int *tracers_per_block, int *thomas_chunk, int *columns_per_block)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    const int nt = c->t.ntr > 0 ? c->t.ntr : 1;
    if (tracers_per_block) *tracers_per_block = std::max(1, std::min(c->step_tpb, nt));
    int tch, cpb;
    fgpu_kd_shape(c, &tch, &cpb);
    if (thomas_chunk) *thomas_chunk = tch;
    if (columns_per_block) *columns_per_block = cpb;
    return FGPU_OK;
}

extern "C" int fgpu_tracer_content(fgpu_context *c, int tr_num, double *content)
{
    int rc = check_ready(c);
    if (rc) return rc;
    if (tr_num < 1 || tr_num > c->t.ntr || !content) return fgpu_fail(FGPU_ERR_ARG, "bad arguments");
    const int np = 1024;
    hipLaunchKernelGGL(k_content_partial, dim3(np), dim3(BLK), 0, c->stream, c->m, c->t.T + c->t.sN * (tr_num - 1), c->d_red);
    LAUNCH_CHECK("k_content_partial");
    hipLaunchKernelGGL(k_content_final, dim3(1), dim3(64), 0, c->stream, c->d_red, np, c->d_red + np);
    LAUNCH_CHECK("k_content_final");
    FGPU_HIP(hipMemcpyAsync(content, c->d_red + np, sizeof(double), hipMemcpyDeviceToHost, c->stream));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    return FGPU_OK;
}
