// Explicit diffusion terms of diff_tracers_ale (oce_ale_tracer.F90:533-793),
// evaluated from T^n before any kernel of the step overwrites it:
//
//   k_hdiff_edge  diff_part_hor_redi (oce_ale_tracer.F90:1526-1689): per edge
//                 and level the flux Kh (grad T + isredi S T_z) through the
//                 edge's dual segments, one thread per (edge, level). The
//                 element gradients (tracer_gradient_elements) and T_z
//                 (tracer_gradient_z, on hnode) are formed on the fly.
//   k_vextra      per column: diff_ver_part_expl_ale (:841-902) for tracers
//                 with i_vert_diff off, and diff_ver_part_redi_expl
//                 (:1439-1522) under Redi.
//
// They write the increments' parts (t.hd, t.ve, t.vr); the node update of
// either step (k_node_update / k_tstar, kD_node) adds them to del_ttf in the
// reference's order: advection, edges in ascending order, explicit vertical,
// explicit Redi, then the T* update. Arithmetic keeps the reference's
// evaluation order, so the sums are bit-identical to a serial run.
#include <algorithm>

#include "fgpu_internal.h"

#define DBLK 256

// T_z at interface k of node n (oce_tracer_mod.F90:192-232): zero at the
// column's top and bottom interfaces and outside the column
__device__ __forceinline__ double grad_z(const DevMesh &m, const double *T, int n, int k)
{
    const int nzmin = m.ulevn[n], nzmax = m.nlevn[n];
    if (k <= nzmin || k >= nzmax) return 0.0;
    const size_t c = (size_t)n * m.L;
    const double dz = 0.5 * (m.hnode[c + k - 2] + m.hnode[c + k - 1]);
    return (T[c + k - 2] - T[c + k - 1]) / dz;
}

// tracer_gradient_elements of element el at level k
__device__ __forceinline__ void grad_el(const DevMesh &m, const double *T, int el, int k, double &gx, double &gy)
{
    const int *en = m.elem_nodes + 3 * (size_t)el;
    const double t0 = T[(size_t)en[0] * m.L + k - 1], t1 = T[(size_t)en[1] * m.L + k - 1],
                 t2 = T[(size_t)en[2] * m.L + k - 1];
    const double *gs = m.gsca + 6 * (size_t)el;
    gx = gs[0] * t0 + gs[1] * t1 + gs[2] * t2;
    gy = gs[3] * t0 + gs[4] * t1 + gs[5] * t2;
}

__global__ void __launch_bounds__(DBLK) k_hdiff_edge(DevMesh m, DevDyn d, DevTracers t, int tr0)
{
    const int tr = tr0 + blockIdx.y;
    const size_t i = (size_t)blockIdx.x * DBLK + threadIdx.x;
    if (i >= (size_t)m.E2 * m.L) return;
    const int e = (int)(i / m.L);
    const int k = (int)(i - (size_t)e * m.L) + 1;
    const int L = m.L;
    const int *ei = m.einfo + 4 * (size_t)e;   // nl1, nu1, nl2, nu2
    const int el1 = m.edge_tri[2 * e], el2 = m.edge_tri[2 * e + 1];
    // cases (A)/(D): el1 alone, (B)/(E): el2 alone, (C): both
    const bool p1 = k >= ei[1] && k <= ei[0];
    const bool p2 = el2 >= 0 && k >= ei[3] && k <= ei[2];
    double *out = t.hd + t.sE * tr + i;
    if (!p1 && !p2) {
        *out = 0.0;
        return;
    }
    const double *T = t.T + t.sN * tr;
    const int n1 = m.edges[2 * e], n2 = m.edges[2 * e + 1];
    const double Kh = (d.Ki[(size_t)n1 * L + k - 1] + d.Ki[(size_t)n2 * L + k - 1]) / 2.0;
    double SxTz = 0.0, SyTz = 0.0;
    const double isredi = d.redi ? 1.0 : 0.0;
    if (d.redi) {
        const double Tz1 = 0.5 * (grad_z(m, T, n1, k) + grad_z(m, T, n1, k + 1));
        const double Tz2 = 0.5 * (grad_z(m, T, n2, k) + grad_z(m, T, n2, k + 1));
        const double *s1 = d.slope + ((size_t)n1 * L + k - 1) * 3, *s2 = d.slope + ((size_t)n2 * L + k - 1) * 3;
        SxTz = (Tz1 * s1[0] + Tz2 * s2[0]) / 2.0;
        SyTz = (Tz1 * s1[1] + Tz2 * s2[1]) / 2.0;
    }
    const double *cd = m.ecd + 4 * (size_t)e;
    const double dX1 = cd[0], dY1 = cd[1];
    double c;
    if (p1 && p2) {
        const double dX2 = cd[2], dY2 = cd[3];
        const double dz = (m.helem[(size_t)el1 * L + k - 1] + m.helem[(size_t)el2 * L + k - 1]) / 2.0;
        double gx1, gy1, gx2, gy2;
        grad_el(m, T, el1, k, gx1, gy1);
        grad_el(m, T, el2, k, gx2, gy2);
        const double Tx = 0.5 * (gx1 + gx2), Ty = 0.5 * (gy1 + gy2);
        const double Fx = Kh * (Tx + SxTz * isredi), Fy = Kh * (Ty + SyTz * isredi);
        c = ((dX2 - dX1) * Fy - (dY2 - dY1) * Fx) * dz;
    } else if (p1) {
        const double dz = m.helem[(size_t)el1 * L + k - 1];
        double Tx, Ty;
        grad_el(m, T, el1, k, Tx, Ty);
        const double Fx = Kh * (Tx + SxTz * isredi), Fy = Kh * (Ty + SyTz * isredi);
        c = (-dX1 * Fy + dY1 * Fx) * dz;
    } else {
        const double dX2 = cd[2], dY2 = cd[3];
        const double dz = m.helem[(size_t)el2 * L + k - 1];
        double Tx, Ty;
        grad_el(m, T, el2, k, Tx, Ty);
        const double Fx = Kh * (Tx + SxTz * isredi), Fy = Kh * (Ty + SyTz * isredi);
        c = (dX2 * Fy - dY2 * Fx) * dz;
    }
    *out = c;
}

// surface flux of the explicit vertical diffusion (oce_ale_tracer.F90:873-885)
__device__ __forceinline__ double expl_surface_flux(int id, const DevDyn &d, int n)
{
    if (id == 1) return -d.heat_flux[n] / 4.2e6;   // vcpw (oce_modules.F90)
    if (id == 2) return d.virtual_salt[n] + d.relax_salt[n] + d.real_salt_flux[n] * d.is_nonlinfs;
    return 0.0;
}

template <int MAXL>
__global__ void __launch_bounds__(64) k_vextra(DevMesh m, DevDyn d, DevTracers t, int tr0, double dt)
{
    const int tr = tr0 + blockIdx.y;
    const int n = blockIdx.x * 64 + threadIdx.x;
    if (n >= m.N2) return;
    const bool expl = !t.ivd[tr];
    const bool redi = d.redi != 0;
    if (!expl && !redi) return;
    const int L = m.L, NL = m.NL;
    const int ul1 = m.ulevn[n], nl1 = m.nlevn[n] - 1;
    const double *T = t.T + t.sN * tr;
    const double *Tn = T + (size_t)n * L;
    const double *ar = m.area + (size_t)n * NL, *av = m.areasvol + (size_t)n * NL;
    double vd[MAXL + 2];
    if (expl) {   // diff_ver_part_expl_ale
        double *ve = t.ve + t.sN * tr + (size_t)n * L;
        const double *kv = d.Kv + (size_t)n * NL, *zb = m.zbar3d + (size_t)n * NL, *Z = m.Z3d + (size_t)n * L;
        vd[ul1] = expl_surface_flux(t.id[tr], d, n);
        for (int nz = ul1 + 1; nz <= nl1; ++nz) {
            const double zinv1 = 1.0 / (Z[nz - 2] - Z[nz - 1]);
            vd[nz] = kv[nz - 1] * (Tn[nz - 2] - Tn[nz - 1]) * zinv1 * ar[nz - 1];
        }
        for (int nz = ul1; nz <= nl1 - 1; ++nz)
            ve[nz - 1] = (vd[nz] - vd[nz + 1]) / (zb[nz - 1] - zb[nz]) * dt / av[nz - 1];
        ve[nl1 - 1] = (vd[nl1] / (zb[nl1 - 1] - zb[nl1])) * dt / av[nl1 - 1];
    }
    if (redi) {   // diff_ver_part_redi_expl
        double *vr = t.vr + t.sN * tr + (size_t)n * L;
        const double *hn = m.hnode + (size_t)n * L;
        const double *ki = d.Ki + (size_t)n * L, *S = d.slope + (size_t)n * L * 3;
        double zbn[MAXL + 2], zn[MAXL + 2], sg[MAXL + 2];
        zbn[nl1 + 1] = m.zbar_n_bot[n];
        zn[nl1] = zbn[nl1 + 1] + hn[nl1 - 1] / 2.0;
        for (int nz = nl1; nz >= ul1 + 1; --nz) {
            zbn[nz] = zbn[nz + 1] + hn[nz - 1];
            zn[nz - 1] = zbn[nz] + hn[nz - 2] / 2.0;
        }
        // node gradients: element gradients weighted by area / 3 / areasvol
        const int *nie = m.nie + (size_t)n * m.kmax;
        const int num = m.nie_num[n];
        for (int nz = ul1; nz <= nl1; ++nz) {
            double Tx = 0.0, Ty = 0.0;
            for (int j = 0; j < num; ++j) {
                const int el = nie[j];
                if (nz <= m.nlev[el] - 1 && nz >= m.ulev[el]) {
                    double gx, gy;
                    grad_el(m, T, el, nz, gx, gy);
                    Tx = Tx + gx * m.elem_area[el];
                    Ty = Ty + gy * m.elem_area[el];
                }
            }
            const double tx = Tx / 3.0 / av[nz - 1], ty = Ty / 3.0 / av[nz - 1];
            sg[nz] = S[(nz - 1) * 3 + 0] * tx + S[(nz - 1) * 3 + 1] * ty;
        }
        vd[ul1] = 0.0;
        for (int nz = ul1 + 1; nz <= nl1; ++nz) {
            double v = (zn[nz - 1] - zbn[nz]) * sg[nz - 1] * ki[nz - 2];
            v = v + (zbn[nz] - zn[nz]) * sg[nz] * ki[nz - 1];
            vd[nz] = v / (zn[nz - 1] - zn[nz]) * ar[nz - 1];
        }
        vd[nl1 + 1] = 0.0;
        for (int nz = ul1; nz <= nl1; ++nz) vr[nz - 1] = (vd[nz] - vd[nz + 1]) * dt / av[nz - 1];
    }
}

// ==============================================================================
// host side
// ==============================================================================
static bool any_explicit_vertical(const fgpu_context *c)
{
    for (int v : c->h_ivd)
        if (!v) return true;
    return false;
}

// buffers of the explicit terms for every tracer slot; before any capture
int fgpu_ensure_diff(fgpu_context *c)
{
    DevTracers &t = c->t;
    const int nt = t.ntr;
    struct A { double **p; size_t n; bool need; } arrs[] = {
        {&t.hd, t.sE * nt, c->d.hdiff != 0}, {&t.ve, t.sN * nt, any_explicit_vertical(c)},
        {&t.vr, t.sN * nt, c->d.redi != 0}};
    for (auto &a : arrs) {
        if (!a.need || *a.p) continue;
        int rc = fgpu_alloc(c, (void **)a.p, a.n * sizeof(double), true);
        if (rc) return rc;
        FGPU_HIP(hipMemset(*a.p, 0, a.n * sizeof(double)));
        c->graph_valid = false;
    }
    return FGPU_OK;
}

// enqueue the explicit terms of tracers tr0..tr0+ntr-1 from the current T
int fgpu_launch_diff(fgpu_context *c, int tr0, int ntr, double dt)
{
    const DevMesh &m = c->m;
    bool ve = false;
    for (int i = tr0; i < tr0 + ntr; ++i) ve = ve || !c->h_ivd[i];
    if (c->d.hdiff) {
        if (!c->t.hd) return fgpu_fail(FGPU_ERR_STATE, "diffusion buffers not allocated");
        ProfScope ps(c, "k_hdiff_edge");
        const size_t n = (size_t)m.E2 * m.L;
        hipLaunchKernelGGL(k_hdiff_edge, dim3((unsigned)((n + DBLK - 1) / DBLK), ntr), dim3(DBLK), 0, c->stream, m, c->d,
                           c->t, tr0);
        FGPU_HIP(hipGetLastError());
    }
    if (ve || c->d.redi) {
        if ((ve && !c->t.ve) || (c->d.redi && !c->t.vr))
            return fgpu_fail(FGPU_ERR_STATE, "diffusion buffers not allocated");
        ProfScope ps(c, "k_vextra");
        const dim3 g((unsigned)((m.N2 + 63) / 64), ntr);
        if (m.L <= 48) hipLaunchKernelGGL(k_vextra<48>, g, dim3(64), 0, c->stream, m, c->d, c->t, tr0, dt);
        else if (m.L <= 72) hipLaunchKernelGGL(k_vextra<72>, g, dim3(64), 0, c->stream, m, c->d, c->t, tr0, dt);
        else hipLaunchKernelGGL(k_vextra<FGPU_MAX_NL>, g, dim3(64), 0, c->stream, m, c->d, c->t, tr0, dt);
        FGPU_HIP(hipGetLastError());
    }
    return FGPU_OK;
}
