// The steps either side of the transport on the device: vert_vel_ale (with
// compute_CFLz and compute_Wvel_split), which turns the horizontal velocity
// into the vertical velocities the transport reads, and compute_ssh_rhs_ale,
// which forms the right-hand side of the SSH solve.
//
// Both are edge->node scatters in the reference (one edge after another,
// oce_ale.F90:1844-1897, 2175-2265). Here each node gathers its edges in
// ascending order, so every node sum adds the same terms in the same order as
// a single-threaded reference run and the results are bit-identical without
// atomics.
#include "fgpu_internal.h"

namespace {

struct VVArgs {
    double dt, maxcfl;
    int which_ale, use_wsplit;
    const double *uv, *hbar, *hbar_old, *water_flux;
    double *w, *w_e, *w_i, *cfl, *hnode_new;
    unsigned long long *cflmax_bits;   // max of CFL_z as the bits of a non-negative double
    int *nan_columns;
};

// One block per node column, one lane per interface k = 1..NL.
//   1. thickness divergence of every layer, edges ascending, el(1) before el(2)
//      (oce_ale.F90:2175-2265)
//   2. cumulative sum from the bottom, division by the interface area
//      (oce_ale.F90:2277-2320)
//   3. zstar: the ssh change spread over the layers above nlevels_nod2D_min,
//      new layer thicknesses, surface water flux (oce_ale.F90:2539-2600)
//   4. CFL_z (oce_ale.F90:2938-2966) and the W split (oce_ale.F90:3035-3047)
template <int LB>
__global__ void __launch_bounds__(LB) k_vert_vel(DevMesh m, VVArgs a)
{
    const int n = blockIdx.x;
    if (n >= m.N2) return;
    __shared__ double sW[LB + 1], sH[LB + 1];
    __shared__ int sBad;
    const int tid = threadIdx.x, k = tid + 1;
    if (tid == 0) sBad = 0;
    const int L = m.L, NL = m.NL;
    const int nzmin = m.ulevn[n], nzmax = m.nlevn[n] - 1;
    double acc = 0.0;
    const int p0 = m.ne_ptr[n], p1 = m.ne_ptr[n + 1];
    for (int p = p0; p < p1; ++p) {
        const int *r = m.ne_rec + 8 * (size_t)p;
        const int q = r[0], e = q >> 1;
        const int nl1 = r[4], nu1 = r[5], nl2 = r[6], nu2 = r[7];
        const int el1 = m.edge_tri[2 * e], el2 = m.edge_tri[2 * e + 1];
        if (k >= nu1 && k <= nl1) {
            const double *u = a.uv + ((size_t)el1 * L + k - 1) * 2;
            const double c = (u[1] * m.ecd[4 * e] - u[0] * m.ecd[4 * e + 1]) * m.helem[(size_t)el1 * L + k - 1];
            acc = (q & 1) ? acc - c : acc + c;
        }
        if (el2 >= 0 && k >= nu2 && k <= nl2) {
            const double *u = a.uv + ((size_t)el2 * L + k - 1) * 2;
            const double c = -((u[1] * m.ecd[4 * e + 2] - u[0] * m.ecd[4 * e + 3]) * m.helem[(size_t)el2 * L + k - 1]);
            acc = (q & 1) ? acc - c : acc + c;
        }
    }
    if (tid < LB) sW[tid] = acc;
    if (tid == 0) sW[LB] = 0.0;
    __syncthreads();
    if (tid == 0)   // serial, bottom up: the reference's association order
        for (int nz = nzmax; nz >= nzmin; --nz) sW[nz - 1] = sW[nz - 1] + sW[nz];
    __syncthreads();
    double w = sW[tid];
    const bool lay = k >= nzmin && k <= nzmax;
    if (lay) w = w / m.area[(size_t)n * NL + k - 1];
    double h = 0.0;
    if (k <= L) h = a.hnode_new[(size_t)n * L + k - 1];
    if (a.which_ale == FGPU_ALE_ZSTAR && nzmin == 1) {
        const int zmax = m.nlevn_min[n] - 1;
        const double dd1 = m.zbar3d[(size_t)n * NL + zmax - 1];
        double dd = m.zbar3d[(size_t)n * NL + nzmin - 1] - dd1;
        dd = (a.hbar[n] - a.hbar_old[n]) / dd;
        const double dddt = dd / a.dt;
        if (k >= nzmin && k <= zmax - 1) {
            const double zk = m.zbar3d[(size_t)n * NL + k - 1];
            w = w - (zk - dd1) * dddt;
            h = m.hnode[(size_t)n * L + k - 1] + (zk - m.zbar3d[(size_t)n * NL + k]) * dd;
            a.hnode_new[(size_t)n * L + k - 1] = h;
        }
        if (k == nzmin) w = w - a.water_flux[n];
    }
    if (lay && w != w) sBad = 1;   // NaN check of oce_ale.F90:2288-2320
    sH[tid] = h;
    __syncthreads();
    if (tid == 0 && sBad) atomicAdd(a.nan_columns, 1);
    // CFL_z(k) = |W(k) dt / h(k-1)| (layer above) + |W(k) dt / h(k)| (layer below)
    double cfl = 0.0;
    if (k >= nzmin + 1 && k <= nzmax + 1) cfl = fabs(w * a.dt / sH[tid - 1]);
    if (lay) cfl = cfl + fabs(w * a.dt / h);
    double we = 0.0, wi = 0.0;
    if (k >= nzmin && k <= nzmax + 1) {
        we = w;
        if (a.use_wsplit && (cfl > a.maxcfl)) {
            const double d = fmax((cfl - a.maxcfl), 0.0) / fmax(a.maxcfl, 1.e-12);
            we = (1.0 / (1.0 + d)) * w;
            wi = (d / (1.0 + d)) * w;
        }
    }
    if (k <= NL) {
        const size_t i = (size_t)n * NL + k - 1;
        a.w[i] = w;
        a.w_e[i] = we;
        a.w_i[i] = wi;
        a.cfl[i] = cfl;
    }
    // block maximum of CFL_z, then one atomic max on its bit pattern (the
    // ordering of non-negative doubles and of their bits agree)
    double cm = k <= NL ? cfl : 0.0;
    for (int off = 32; off > 0; off >>= 1) cm = fmax(cm, __shfl_xor(cm, off));
    __syncthreads();
    if ((tid & 63) == 0) sW[tid >> 6] = cm;
    __syncthreads();
    if (tid == 0) {
        for (int i = 1; i < LB / 64; ++i) cm = fmax(cm, sW[i]);
        atomicMax(a.cflmax_bits, (unsigned long long)__double_as_longlong(cm));
    }
}

// Depth-integrated transport of alpha (u + u_rhs) through the two dual faces
// of edge e, levels in order (oce_ale.F90:1844-1880); one thread per edge.
__global__ void k_ssh_rhs_edge(DevMesh m, const double *uv, const double *uvr, double alpha, double *ce)
{
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= m.E2) return;
    const int L = m.L;
    const int el1 = m.edge_tri[2 * e], el2 = m.edge_tri[2 * e + 1];
    double c1 = 0.0, c2 = 0.0;
    {
        const double dX = m.ecd[4 * e], dY = m.ecd[4 * e + 1];
        const int nu = m.ulev[el1], nlv = m.nlev[el1] - 1;
        for (int nz = nu; nz <= nlv; ++nz) {
            const size_t i = ((size_t)el1 * L + nz - 1) * 2;
            c1 = c1 + alpha * ((uv[i + 1] + uvr[i + 1]) * dX - (uv[i] + uvr[i]) * dY) * m.helem[(size_t)el1 * L + nz - 1];
        }
    }
    if (el2 >= 0) {
        const double dX = m.ecd[4 * e + 2], dY = m.ecd[4 * e + 3];
        const int nu = m.ulev[el2], nlv = m.nlev[el2] - 1;
        for (int nz = nu; nz <= nlv; ++nz) {
            const size_t i = ((size_t)el2 * L + nz - 1) * 2;
            c2 = c2 - alpha * ((uv[i + 1] + uvr[i + 1]) * dX - (uv[i] + uvr[i]) * dY) * m.helem[(size_t)el2 * L + nz - 1];
        }
    }
    ce[e] = c1 + c2;
}

// Node sums over the edges in ascending order (oce_ale.F90:1889-1894), then
// the free-surface terms (oce_ale.F90:1913-1934); one thread per node.
__global__ void k_ssh_rhs_node(DevMesh m, const double *ce, const double *old, const double *wf, double alpha,
                               int which_ale, int cavity_fw2press, double *rhs)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.N2) return;
    double s = 0.0;
    for (int p = m.ne_ptr[n]; p < m.ne_ptr[n + 1]; ++p) {
        const int q = m.ne_idx[p];
        s = (q & 1) ? s - ce[q >> 1] : s + ce[q >> 1];
    }
    if (which_ale != FGPU_ALE_LINFS) {
        const int nzmin = m.ulevn[n];
        const double av = m.areasvol[(size_t)n * m.NL + nzmin - 1];
        if (nzmin > 1) {
            if (cavity_fw2press) s = s - alpha * wf[n] * av;
        } else {
            s = s - alpha * wf[n] * av + (1.0 - alpha) * old[n];
        }
    } else {
        s = s + (1.0 - alpha) * old[n];
    }
    rhs[n] = s;
}

int ensure_buf(fgpu_context *c, double **p, size_t n)
{
    if (*p) return FGPU_OK;
    return fgpu_alloc(c, (void **)p, sizeof(double) * n);
}

int ensure_const_buf(fgpu_context *c, const double **p, size_t n)
{
    if (*p) return FGPU_OK;
    double *q = nullptr;
    int rc = fgpu_alloc(c, (void **)&q, sizeof(double) * n);
    if (rc) return rc;
    FGPU_HIP(hipMemset(q, 0, sizeof(double) * n));
    *p = q;
    return FGPU_OK;
}

int upload_to(fgpu_context *c, double **dst, const double *src, size_t n)
{
    int rc = ensure_buf(c, dst, n);
    if (rc) return rc;
    FGPU_HIP(hipMemcpyAsync(*dst, src, sizeof(double) * n, hipMemcpyHostToDevice, c->stream));
    return FGPU_OK;
}

int check_dyn(fgpu_context *c)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    if (!c->vert_ready) return fgpu_fail(FGPU_ERR_STATE, "vertical geometry not set (fgpu_set_vertical)");
    if (!c->d.uv) return fgpu_fail(FGPU_ERR_STATE, "horizontal velocity not set (fgpu_set_uv or fgpu_set_velocity)");
    FGPU_HIP(hipSetDevice(c->device));
    return FGPU_OK;
}

}  // namespace

extern "C" int fgpu_set_uv(fgpu_context *c, const double *uv)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    if (!uv) return fgpu_fail(FGPU_ERR_ARG, "null array");
    FGPU_HIP(hipSetDevice(c->device));
    const size_t n = (size_t)c->F2 * (c->nl - 1) * 2;
    if (!c->d.uv) {
        double *p = nullptr;
        int rc = fgpu_alloc(c, (void **)&p, sizeof(double) * n);
        if (rc) return rc;
        c->d.uv = p;
    }
    FGPU_HIP(hipMemcpy((void *)c->d.uv, uv, sizeof(double) * n, hipMemcpyHostToDevice));
    return FGPU_OK;
}

extern "C" int fgpu_vert_vel_ale(fgpu_context *c, double dt, int which_ale, const double *hbar,
                                 const double *hbar_old, int use_wsplit, double wsplit_maxcfl, double *cflmax,
                                 int *nan_columns)
{
    int rc = check_dyn(c);
    if (rc) return rc;
    if (which_ale == FGPU_ALE_ZLEVEL)
        return fgpu_fail(FGPU_ERR_ARG, "which_ALE='zlevel' (local zstar fallback) is not provided");
    if (which_ale != FGPU_ALE_LINFS && which_ale != FGPU_ALE_ZSTAR) return fgpu_fail(FGPU_ERR_ARG, "unknown which_ALE");
    if (!(dt > 0.0)) return fgpu_fail(FGPU_ERR_ARG, "dt must be > 0");
    if (c->nl > 128) return fgpu_fail(FGPU_ERR_ARG, "at most 128 interfaces per column");
    const size_t N2 = c->N2, NL = c->nl;
    VVArgs a{};
    a.dt = dt;
    a.maxcfl = wsplit_maxcfl;
    a.which_ale = which_ale;
    a.use_wsplit = use_wsplit ? 1 : 0;
    a.uv = c->d.uv;
    if (which_ale == FGPU_ALE_ZSTAR) {
        if (!hbar || !hbar_old) return fgpu_fail(FGPU_ERR_ARG, "zstar needs hbar and hbar_old");
        if (!c->forcing_ready) return fgpu_fail(FGPU_ERR_STATE, "zstar needs the surface water flux (fgpu_set_surface_forcing)");
        if ((rc = upload_to(c, &c->vv_hbar, hbar, N2))) return rc;
        if ((rc = upload_to(c, &c->vv_hbar_old, hbar_old, N2))) return rc;
        a.hbar = c->vv_hbar;
        a.hbar_old = c->vv_hbar_old;
        a.water_flux = c->d.water_flux;
    }
    if ((rc = ensure_const_buf(c, &c->d.w, N2 * NL))) return rc;
    if ((rc = ensure_const_buf(c, &c->d.w_e, N2 * NL))) return rc;
    if ((rc = ensure_const_buf(c, &c->d.w_i, N2 * NL))) return rc;
    if ((rc = ensure_buf(c, &c->vv_cfl, N2 * NL))) return rc;
    if (!c->vv_flags) {
        void *p = nullptr;
        if ((rc = fgpu_alloc(c, &p, 64))) return rc;
        c->vv_flags = p;
    }
    a.w = (double *)c->d.w;
    a.w_e = (double *)c->d.w_e;
    a.w_i = (double *)c->d.w_i;
    a.cfl = c->vv_cfl;
    a.hnode_new = (double *)c->m.hnode_new;
    a.cflmax_bits = (unsigned long long *)c->vv_flags;
    a.nan_columns = (int *)((char *)c->vv_flags + 8);
    FGPU_HIP(hipMemsetAsync(c->vv_flags, 0, 16, c->stream));
    {
        ProfScope ps(c, "k_vert_vel");
        if (NL <= 64) hipLaunchKernelGGL(k_vert_vel<64>, dim3(c->N2), dim3(64), 0, c->stream, c->m, a);
        else hipLaunchKernelGGL(k_vert_vel<128>, dim3(c->N2), dim3(128), 0, c->stream, c->m, a);
        FGPU_HIP(hipGetLastError());
    }
    if (which_ale == FGPU_ALE_ZSTAR && (rc = fgpu_compute_zn(c))) return rc;   // Z levels of the implicit diffusion
    unsigned long long flags[2] = {0ull, 0ull};
    FGPU_HIP(hipMemcpyAsync(flags, c->vv_flags, 16, hipMemcpyDeviceToHost, c->stream));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    if (c->prof && (rc = fgpu_prof_flush(c))) return rc;
    if (cflmax) *cflmax = __builtin_bit_cast(double, flags[0]);
    if (nan_columns) *nan_columns = (int)(flags[1] & 0xffffffffull);
    c->wsplit_active = use_wsplit != 0;
    c->vel_ready = true;
    return FGPU_OK;
}

extern "C" int fgpu_vert_vel_download(fgpu_context *c, double *w, double *w_e, double *w_i, double *cfl_z,
                                      double *hnode_new)
{
    int rc = check_dyn(c);
    if (rc) return rc;
    const size_t N2 = c->N2, NL = c->nl;
    FGPU_HIP(hipStreamSynchronize(c->stream));
    struct { double *dst; const double *src; size_t n; } cp[] = {
        {w, c->d.w, N2 * NL}, {w_e, c->d.w_e, N2 * NL}, {w_i, c->d.w_i, N2 * NL}, {cfl_z, c->vv_cfl, N2 * NL},
        {hnode_new, c->m.hnode_new, N2 * (NL - 1)}};
    for (auto &x : cp) {
        if (!x.dst) continue;
        if (!x.src) return fgpu_fail(FGPU_ERR_STATE, "nothing computed yet (fgpu_vert_vel_ale)");
        FGPU_HIP(hipMemcpy(x.dst, x.src, sizeof(double) * x.n, hipMemcpyDeviceToHost));
    }
    return FGPU_OK;
}

extern "C" int fgpu_compute_ssh_rhs_ale(fgpu_context *c, const double *uv_rhs, const double *ssh_rhs_old,
                                        double alpha, int which_ale, int use_cavity_fw2press, double *ssh_rhs)
{
    int rc = check_dyn(c);
    if (rc) return rc;
    if (!uv_rhs || !ssh_rhs_old) return fgpu_fail(FGPU_ERR_ARG, "null array");
    if (which_ale != FGPU_ALE_LINFS && which_ale != FGPU_ALE_ZLEVEL && which_ale != FGPU_ALE_ZSTAR)
        return fgpu_fail(FGPU_ERR_ARG, "unknown which_ALE");
    if (which_ale != FGPU_ALE_LINFS && !c->forcing_ready)
        return fgpu_fail(FGPU_ERR_STATE, "the full free surface needs the water flux (fgpu_set_surface_forcing)");
    const size_t N2 = c->N2, E2 = c->E2, F2 = c->F2, L = c->nl - 1;
    if ((rc = upload_to(c, &c->vv_uvr, uv_rhs, F2 * L * 2))) return rc;
    if ((rc = upload_to(c, &c->vv_old, ssh_rhs_old, N2))) return rc;
    if ((rc = ensure_buf(c, &c->vv_ce, E2))) return rc;
    if ((rc = ensure_buf(c, &c->vv_rhs, N2))) return rc;
    const double *wf = c->d.water_flux ? c->d.water_flux : c->vv_old;   // unread under linfs
    {
        ProfScope ps(c, "k_ssh_rhs_edge");
        hipLaunchKernelGGL(k_ssh_rhs_edge, dim3((c->E2 + 255) / 256), dim3(256), 0, c->stream, c->m, c->d.uv,
                           c->vv_uvr, alpha, c->vv_ce);
        FGPU_HIP(hipGetLastError());
    }
    {
        ProfScope ps(c, "k_ssh_rhs_node");
        hipLaunchKernelGGL(k_ssh_rhs_node, dim3((c->N2 + 255) / 256), dim3(256), 0, c->stream, c->m, c->vv_ce,
                           c->vv_old, wf, alpha, which_ale, use_cavity_fw2press ? 1 : 0, c->vv_rhs);
        FGPU_HIP(hipGetLastError());
    }
    // the SSH solve's right-hand side when its system is set up
    if (c->ssh_ready && c->ssh_n == c->N2)
        FGPU_HIP(hipMemcpyAsync(c->ssh_rhs, c->vv_rhs, sizeof(double) * N2, hipMemcpyDeviceToDevice, c->stream));
    if (ssh_rhs) FGPU_HIP(hipMemcpyAsync(ssh_rhs, c->vv_rhs, sizeof(double) * N2, hipMemcpyDeviceToHost, c->stream));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    if (c->prof && (rc = fgpu_prof_flush(c))) return rc;
    return FGPU_OK;
}
