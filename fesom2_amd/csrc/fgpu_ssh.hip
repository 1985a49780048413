// SSH elliptic operator and preconditioned conjugate gradients for gfx950.
//
// * init_stiff_mat_ale (oce_ale.F90:1393-1582): one thread per matrix row
//   gathers the edge/element contributions of its row in ascending edge
//   order (the reference's serial order), so the assembly is deterministic
//   and free of atomics.
// * ssh_solve_preconditioner (solver.F90:31-117): one thread per row.
// * ssh_solve_cg (solver.F90:120-385): one persistent cooperative launch holds
//   the whole iteration on the device. Rows are owned by a fixed grid-stride
//   assignment; each reduction writes one partial per block, the grid
//   synchronises, and every block then adds the partials in the same fixed
//   order, so all blocks see the same bit pattern and take the same branch
//   (convergence, breakdown) without a host round trip. Iterates, residual
//   and dot products are FP64.
#include <hip/hip_cooperative_groups.h>

#include <cmath>
#include <cstring>

#include "fgpu_internal.h"

namespace cg = cooperative_groups;

#define SBLK 256

// ------------------------------------------------------------------------------
// assembly
// ------------------------------------------------------------------------------
__global__ void __launch_bounds__(SBLK) k_stiff_rows(DevMesh m, const int *rowptr, const int *colind, double *val,
                                                     double dt, double factor)
{
    const int r = blockIdx.x * SBLK + threadIdx.x;
    if (r >= m.N2) return;
    const int p0 = rowptr[r], p1 = rowptr[r + 1];
    for (int p = p0; p < p1; ++p) val[p] = 0.0;
    for (int pp = m.ne_ptr[r]; pp < m.ne_ptr[r + 1]; ++pp) {
        const int q = m.ne_idx[pp];
        const int ed = q >> 1, side = q & 1;
        for (int i = 0; i < 2; ++i) {
            const int el = m.edge_tri[2 * ed + i];
            if (el < 0) continue;
            const int *en = m.elem_nodes + 3 * (size_t)el;
            const double *gs = m.gsca + 6 * (size_t)el;
            const double hh = m.zbar_e_bot[el] - m.zbar_e_srf[el];
            const double cx = m.ecd[4 * ed + 2 * i];
            const double cy = m.ecd[4 * ed + 2 * i + 1];
            double fy[3];
            for (int j = 0; j < 3; ++j) fy[j] = hh * (gs[j] * cy - gs[3 + j] * cx);
            if (i == 1)
                for (int j = 0; j < 3; ++j) fy[j] = -fy[j];
            for (int j = 0; j < 3; ++j) {
                int pos = -1;
                for (int p = p0; p < p1; ++p)
                    if (colind[p] == en[j]) { pos = p; break; }
                if (pos < 0) continue;   // cannot happen on a valid mesh
                if (side == 0) val[pos] = val[pos] + fy[j] * factor;
                else val[pos] = val[pos] - fy[j] * factor;
            }
        }
    }
    if (m.ulevn[r] <= 1) val[p0] = val[p0] + m.areasvol[(size_t)r * m.NL + m.ulevn[r] - 1] / dt;
}

__global__ void __launch_bounds__(SBLK) k_precond(int n, const int *rowptr, const int *colind, const double *val,
                                                  int variant, double *pr, double *dinv)
{
    const int r = blockIdx.x * SBLK + threadIdx.x;
    if (r >= n) return;
    const int off = rowptr[r];
    const int nend = rowptr[r + 1] - off;
    pr[off] = 1.0 / val[off];
    dinv[r] = pr[off];
    for (int j = 1; j < nend; ++j) {
        const int node = colind[off + j];
        const double dnode = val[rowptr[node]];
        if (variant == 0) pr[off + j] = -0.5 * (val[off + j] / val[off]) / (val[off] + dnode);
        else if (variant == 1) pr[off + j] = -val[off + j] / (val[off] * dnode);
        else pr[off + j] = 0.0;
    }
}

// ------------------------------------------------------------------------------
// CG
// ------------------------------------------------------------------------------
struct CGArgs {
    int n, variant, maxiter;
    const int *rowptr, *colind;
    const double *val, *pr, *dinv;
    double *x;
    const double *rhs;
    double *r, *z, *p, *q;
    double *part;   // 6 slots x nblocks
    double *out;    // iters, converged, nbreakdown, nnegrz, resid, rtol
    double soltol;
};

__device__ __forceinline__ double row_dot(const int *rowptr, const int *colind, const double *v, const double *x, int r)
{
    double s = 0.0;
    for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) s = s + v[p] * x[colind[p]];
    return s;
}

// deterministic grid-wide sum of one or two values
__device__ void grid_sum2(double a, double b, double *part, int slot, cg::grid_group &g, double &sa, double &sb)
{
    __shared__ double sh[2][SBLK];
    const int tid = threadIdx.x, nb = gridDim.x;
    sh[0][tid] = a;
    sh[1][tid] = b;
    __syncthreads();
    for (int w = SBLK / 2; w > 0; w >>= 1) {
        if (tid < w) {
            sh[0][tid] += sh[0][tid + w];
            sh[1][tid] += sh[1][tid + w];
        }
        __syncthreads();
    }
    if (tid == 0) {
        part[(size_t)slot * nb + blockIdx.x] = sh[0][0];
        part[(size_t)(slot + 1) * nb + blockIdx.x] = sh[1][0];
    }
    __threadfence();
    g.sync();
    double s0 = 0.0, s1 = 0.0;
    for (int j = tid; j < nb; j += SBLK) {
        s0 += part[(size_t)slot * nb + j];
        s1 += part[(size_t)(slot + 1) * nb + j];
    }
    __syncthreads();
    sh[0][tid] = s0;
    sh[1][tid] = s1;
    __syncthreads();
    for (int w = SBLK / 2; w > 0; w >>= 1) {
        if (tid < w) {
            sh[0][tid] += sh[0][tid + w];
            sh[1][tid] += sh[1][tid + w];
        }
        __syncthreads();
    }
    sa = sh[0][0];
    sb = sh[1][0];
    __syncthreads();
}

__device__ __forceinline__ void apply_prec(const CGArgs &A, int row)
{
    if (A.variant == 2) A.z[row] = A.dinv[row] * A.r[row];
    else A.z[row] = row_dot(A.rowptr, A.colind, A.pr, A.r, row);
}

__global__ void __launch_bounds__(SBLK) k_cg(CGArgs A)
{
    cg::grid_group g = cg::this_grid();
    const int gsize = gridDim.x * SBLK;
    const int gtid = blockIdx.x * SBLK + threadIdx.x;
    const int n = A.n;
    const bool sparse_m = A.variant != 2;
    double a0 = 0.0, dum;
    for (int row = gtid; row < n; row += gsize) a0 += A.rhs[row] * A.rhs[row];
    double s_old;
    grid_sum2(a0, 0.0, A.part, 0, g, s_old, dum);
    const double rtol = A.soltol * sqrt(s_old / (double)n);
    for (int row = gtid; row < n; row += gsize)
        A.r[row] = A.rhs[row] - row_dot(A.rowptr, A.colind, A.val, A.x, row);
    if (sparse_m) { __threadfence(); g.sync(); }
    a0 = 0.0;
    for (int row = gtid; row < n; row += gsize) {
        apply_prec(A, row);
        A.p[row] = A.z[row];
        a0 += A.r[row] * A.z[row];
    }
    grid_sum2(a0, 0.0, A.part, 2, g, s_old, dum);
    int iter, converged = 0, nbreak = 0, nneg = 0;
    double rr = 0.0;
    for (iter = 1; iter <= A.maxiter; ++iter) {
        __threadfence();
        g.sync();   // p complete before the matrix-vector product reads it
        a0 = 0.0;
        for (int row = gtid; row < n; row += gsize) {
            const double qq = row_dot(A.rowptr, A.colind, A.val, A.p, row);
            A.q[row] = qq;
            a0 += A.p[row] * qq;
        }
        double s_aux;
        grid_sum2(a0, 0.0, A.part, 2, g, s_aux, dum);
        if (s_aux <= 0.0 || s_aux != s_aux) { nbreak++; break; }
        const double al = s_old / s_aux;
        for (int row = gtid; row < n; row += gsize) {
            A.x[row] = A.x[row] + al * A.p[row];
            A.r[row] = A.r[row] - al * A.q[row];
        }
        if (sparse_m) { __threadfence(); g.sync(); }
        double b0 = 0.0;
        a0 = 0.0;
        for (int row = gtid; row < n; row += gsize) {
            apply_prec(A, row);
            a0 += A.r[row] * A.z[row];
            b0 += A.r[row] * A.r[row];
        }
        double rz;
        grid_sum2(a0, b0, A.part, 4, g, rz, rr);
        if (sqrt(rr / (double)n) < rtol) { converged = 1; break; }
        if (rz < 0.0) nneg++;
        if (s_old == 0.0) { nbreak++; break; }
        const double be = rz / s_old;
        s_old = rz;
        for (int row = gtid; row < n; row += gsize) A.p[row] = A.z[row] + be * A.p[row];
    }
    if (gtid == 0) {
        A.out[0] = (double)(iter < A.maxiter ? iter : A.maxiter);
        A.out[1] = converged;
        A.out[2] = nbreak;
        A.out[3] = nneg;
        A.out[4] = sqrt(rr / (double)n);
        A.out[5] = rtol;
    }
}

__global__ void k_copy(double *dst, const double *src, int n)
{
    const int i = blockIdx.x * SBLK + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

// ==============================================================================
// host side
// ==============================================================================
static int ssh_alloc_vectors(fgpu_context *c)
{
    const size_t n = c->ssh_n;
    int rc;
    double **vs[] = {&c->ssh_x, &c->ssh_x0, &c->ssh_rhs, &c->ssh_r, &c->ssh_z, &c->ssh_p, &c->ssh_q, &c->ssh_diag_inv};
    for (double **v : vs)
        if (!*v && (rc = fgpu_alloc(c, (void **)v, n * sizeof(double)))) return rc;
    int nb = 0, dev = c->device, ncu = 0;
    FGPU_HIP(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
    FGPU_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_cg, SBLK, 0));
    if (nb < 1) return fgpu_fail(FGPU_ERR_HIP, "k_cg cannot be resident");
    int blocks = ncu * std::min(nb, 2);
    int need = (int)((n + SBLK - 1) / SBLK);
    if (blocks > need) blocks = need;
    if (blocks < 1) blocks = 1;
    c->ssh_blocks = blocks;
    if (!c->ssh_part && (rc = fgpu_alloc(c, (void **)&c->ssh_part, sizeof(double) * 6 * (size_t)blocks))) return rc;
    if (!c->ssh_out && (rc = fgpu_alloc(c, (void **)&c->ssh_out, sizeof(double) * 8))) return rc;
    return FGPU_OK;
}

static int ssh_pattern(fgpu_context *c)
{
    // oce_ale.F90:1420-1466: diagonal first, then neighbours in edge order
    const int N2 = c->N2;
    std::vector<int> num(N2, 1);
    std::vector<std::vector<int>> pos(N2);
    for (int n = 0; n < N2; ++n) pos[n].push_back(n);
    for (int e = 0; e < c->E2; ++e) {
        int n1 = c->h_edges[2 * e], n2 = c->h_edges[2 * e + 1];
        pos[n1].push_back(n2);
        pos[n2].push_back(n1);
    }
    c->h_ssh_rowptr.assign(N2 + 1, 0);
    for (int n = 0; n < N2; ++n) c->h_ssh_rowptr[n + 1] = c->h_ssh_rowptr[n] + (int)pos[n].size();
    c->h_ssh_colind.resize(c->h_ssh_rowptr[N2]);
    for (int n = 0; n < N2; ++n)
        std::copy(pos[n].begin(), pos[n].end(), c->h_ssh_colind.begin() + c->h_ssh_rowptr[n]);
    return FGPU_OK;
}

static int ssh_alloc_matrix(fgpu_context *c, int n, int nnz)
{
    int rc;
    if (c->ssh_ready && (c->ssh_n != n || c->ssh_nnz != nnz))
        return fgpu_fail(FGPU_ERR_STATE, "SSH matrix shape changed");
    c->ssh_n = n;
    c->ssh_nnz = nnz;
    if (!c->ssh_rowptr && (rc = fgpu_alloc(c, (void **)&c->ssh_rowptr, sizeof(int) * (n + 1)))) return rc;
    if (!c->ssh_colind && (rc = fgpu_alloc(c, (void **)&c->ssh_colind, sizeof(int) * nnz))) return rc;
    if (!c->ssh_val && (rc = fgpu_alloc(c, (void **)&c->ssh_val, sizeof(double) * nnz))) return rc;
    if (!c->ssh_pr && (rc = fgpu_alloc(c, (void **)&c->ssh_pr, sizeof(double) * nnz))) return rc;
    return ssh_alloc_vectors(c);
}

extern "C" int fgpu_init_stiff_mat_ale(fgpu_context *c, double dt, double g, double alpha, double theta)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->mesh_ready || !c->vert_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh/vertical geometry not set");
    FGPU_HIP(hipSetDevice(c->device));
    int rc;
    if ((rc = ssh_pattern(c))) return rc;
    const int n = c->N2, nnz = (int)c->h_ssh_colind.size();
    if ((rc = ssh_alloc_matrix(c, n, nnz))) return rc;
    FGPU_HIP(hipMemcpy(c->ssh_rowptr, c->h_ssh_rowptr.data(), sizeof(int) * (n + 1), hipMemcpyHostToDevice));
    FGPU_HIP(hipMemcpy(c->ssh_colind, c->h_ssh_colind.data(), sizeof(int) * nnz, hipMemcpyHostToDevice));
    const double factor = g * dt * alpha * theta;
    hipLaunchKernelGGL(k_stiff_rows, dim3((n + SBLK - 1) / SBLK), dim3(SBLK), 0, c->stream, c->m, c->ssh_rowptr,
                       c->ssh_colind, c->ssh_val, dt, factor);
    FGPU_HIP(hipGetLastError());
    FGPU_HIP(hipStreamSynchronize(c->stream));
    c->ssh_ready = true;
    c->ssh_pr_ready = false;
    return FGPU_OK;
}

extern "C" int fgpu_ssh_set_matrix(fgpu_context *c, int dim, int nza, const int *rowptr, const int *colind,
                                   const double *values)
{
    if (!c || !rowptr || !colind || !values) return fgpu_fail(FGPU_ERR_ARG, "null argument");
    if (dim < 1 || nza < dim) return fgpu_fail(FGPU_ERR_ARG, "bad matrix size");
    if (rowptr[0] != 1 || rowptr[dim] - 1 != nza) return fgpu_fail(FGPU_ERR_ARG, "rowptr must be 1-based CSR");
    FGPU_HIP(hipSetDevice(c->device));
    std::vector<int> rp(dim + 1), ci(nza);
    for (int i = 0; i <= dim; ++i) rp[i] = rowptr[i] - 1;
    for (int r = 0; r < dim; ++r) {
        if (rp[r + 1] <= rp[r]) return fgpu_fail(FGPU_ERR_ARG, "empty matrix row");
        if (colind[rp[r]] - 1 != r) return fgpu_fail(FGPU_ERR_ARG, "diagonal must come first in every row");
    }
    for (int i = 0; i < nza; ++i) {
        ci[i] = colind[i] - 1;
        if (ci[i] < 0 || ci[i] >= dim) return fgpu_fail(FGPU_ERR_ARG, "colind out of range");
    }
    int rc;
    if ((rc = ssh_alloc_matrix(c, dim, nza))) return rc;
    FGPU_HIP(hipMemcpy(c->ssh_rowptr, rp.data(), sizeof(int) * (dim + 1), hipMemcpyHostToDevice));
    FGPU_HIP(hipMemcpy(c->ssh_colind, ci.data(), sizeof(int) * nza, hipMemcpyHostToDevice));
    FGPU_HIP(hipMemcpy(c->ssh_val, values, sizeof(double) * nza, hipMemcpyHostToDevice));
    c->h_ssh_rowptr = rp;
    c->h_ssh_colind = ci;
    c->ssh_ready = true;
    c->ssh_pr_ready = false;
    return FGPU_OK;
}

extern "C" int fgpu_ssh_get_matrix(fgpu_context *c, int *rowptr, int *colind, double *values, double *pr_values,
                                   int *nza)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->ssh_ready) return fgpu_fail(FGPU_ERR_STATE, "SSH matrix not set");
    FGPU_HIP(hipSetDevice(c->device));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    if (nza) *nza = c->ssh_nnz;
    if (rowptr)
        for (int i = 0; i <= c->ssh_n; ++i) rowptr[i] = c->h_ssh_rowptr[i] + 1;
    if (colind)
        for (int i = 0; i < c->ssh_nnz; ++i) colind[i] = c->h_ssh_colind[i] + 1;
    if (values) FGPU_HIP(hipMemcpy(values, c->ssh_val, sizeof(double) * c->ssh_nnz, hipMemcpyDeviceToHost));
    if (pr_values) {
        if (!c->ssh_pr_ready) return fgpu_fail(FGPU_ERR_STATE, "preconditioner not built");
        FGPU_HIP(hipMemcpy(pr_values, c->ssh_pr, sizeof(double) * c->ssh_nnz, hipMemcpyDeviceToHost));
    }
    return FGPU_OK;
}

extern "C" int fgpu_ssh_solve_preconditioner(fgpu_context *c, int variant)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->ssh_ready) return fgpu_fail(FGPU_ERR_STATE, "SSH matrix not set");
    if (variant < 0 || variant > 2) return fgpu_fail(FGPU_ERR_ARG, "precond_variant must be 0, 1 or 2");
    FGPU_HIP(hipSetDevice(c->device));
    hipLaunchKernelGGL(k_precond, dim3((c->ssh_n + SBLK - 1) / SBLK), dim3(SBLK), 0, c->stream, c->ssh_n,
                       c->ssh_rowptr, c->ssh_colind, c->ssh_val, variant, c->ssh_pr, c->ssh_diag_inv);
    FGPU_HIP(hipGetLastError());
    FGPU_HIP(hipStreamSynchronize(c->stream));
    c->ssh_variant = variant;
    c->ssh_pr_ready = true;
    return FGPU_OK;
}

static int launch_cg(fgpu_context *c, double soltol, int maxiter)
{
    CGArgs A;
    A.n = c->ssh_n;
    A.variant = c->ssh_variant;
    A.maxiter = maxiter;
    A.rowptr = c->ssh_rowptr;
    A.colind = c->ssh_colind;
    A.val = c->ssh_val;
    A.pr = c->ssh_pr;
    A.dinv = c->ssh_diag_inv;
    A.x = c->ssh_x;
    A.rhs = c->ssh_rhs;
    A.r = c->ssh_r;
    A.z = c->ssh_z;
    A.p = c->ssh_p;
    A.q = c->ssh_q;
    A.part = c->ssh_part;
    A.out = c->ssh_out;
    A.soltol = soltol;
    void *args[] = {&A};
    FGPU_HIP(hipLaunchCooperativeKernel((const void *)k_cg, dim3(c->ssh_blocks), dim3(SBLK), args, 0, c->stream));
    return FGPU_OK;
}

static int read_stats(fgpu_context *c, fgpu_cg_stats *st, double ms)
{
    double o[6];
    FGPU_HIP(hipMemcpy(o, c->ssh_out, sizeof(o), hipMemcpyDeviceToHost));
    if (st) {
        st->iters_last = (int)o[0];
        st->converged = (int)o[1];
        st->nbreakdown = (int)o[2];
        st->nnegrz = (int)o[3];
        st->resid_last = o[4];
        st->rtol_last = o[5];
        st->ms = ms;
    }
    return FGPU_OK;
}

static int cg_ready(fgpu_context *c, double soltol, int maxiter)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->ssh_ready) return fgpu_fail(FGPU_ERR_STATE, "SSH matrix not set");
    if (!c->ssh_pr_ready) return fgpu_fail(FGPU_ERR_STATE, "preconditioner not built (fgpu_ssh_solve_preconditioner)");
    if (maxiter < 1 || !(soltol > 0.0)) return fgpu_fail(FGPU_ERR_ARG, "soltol must be > 0 and maxiter >= 1");
    FGPU_HIP(hipSetDevice(c->device));
    return FGPU_OK;
}

extern "C" int fgpu_ssh_solve_cg(fgpu_context *c, double *x, const double *rhs, double soltol, int maxiter,
                                 fgpu_cg_stats *stats)
{
    int rc = cg_ready(c, soltol, maxiter);
    if (rc) return rc;
    if (!x || !rhs) return fgpu_fail(FGPU_ERR_ARG, "null vector");
    const size_t nb = sizeof(double) * c->ssh_n;
    FGPU_HIP(hipMemcpyAsync(c->ssh_x, x, nb, hipMemcpyHostToDevice, c->stream));
    FGPU_HIP(hipMemcpyAsync(c->ssh_rhs, rhs, nb, hipMemcpyHostToDevice, c->stream));
    hipEvent_t e0, e1;
    FGPU_HIP(hipEventCreate(&e0));
    FGPU_HIP(hipEventCreate(&e1));
    FGPU_HIP(hipEventRecord(e0, c->stream));
    rc = launch_cg(c, soltol, maxiter);
    FGPU_HIP(hipEventRecord(e1, c->stream));
    FGPU_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    FGPU_HIP(hipMemcpy(x, c->ssh_x, nb, hipMemcpyDeviceToHost));
    return read_stats(c, stats, ms);
}

extern "C" int fgpu_ssh_upload(fgpu_context *c, const double *x0, const double *rhs)
{
    if (!c || !x0 || !rhs) return fgpu_fail(FGPU_ERR_ARG, "null argument");
    if (!c->ssh_ready) return fgpu_fail(FGPU_ERR_STATE, "SSH matrix not set");
    FGPU_HIP(hipSetDevice(c->device));
    const size_t nb = sizeof(double) * c->ssh_n;
    FGPU_HIP(hipMemcpy(c->ssh_x0, x0, nb, hipMemcpyHostToDevice));
    FGPU_HIP(hipMemcpy(c->ssh_rhs, rhs, nb, hipMemcpyHostToDevice));
    return FGPU_OK;
}

extern "C" int fgpu_ssh_solve_resident(fgpu_context *c, double soltol, int maxiter, int nrep, fgpu_cg_stats *stats)
{
    int rc = cg_ready(c, soltol, maxiter);
    if (rc) return rc;
    if (nrep < 1) return fgpu_fail(FGPU_ERR_ARG, "nrep < 1");
    hipEvent_t e0, e1;
    FGPU_HIP(hipEventCreate(&e0));
    FGPU_HIP(hipEventCreate(&e1));
    float total = 0.f;
    for (int i = 0; i < nrep && !rc; ++i) {
        hipLaunchKernelGGL(k_copy, dim3((c->ssh_n + SBLK - 1) / SBLK), dim3(SBLK), 0, c->stream, c->ssh_x, c->ssh_x0,
                           c->ssh_n);
        FGPU_HIP(hipGetLastError());
        FGPU_HIP(hipEventRecord(e0, c->stream));
        rc = launch_cg(c, soltol, maxiter);
        FGPU_HIP(hipEventRecord(e1, c->stream));
        FGPU_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        FGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
        total += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    return read_stats(c, stats, total / nrep);
}

extern "C" int fgpu_ssh_download(fgpu_context *c, double *x)
{
    if (!c || !x) return fgpu_fail(FGPU_ERR_ARG, "null argument");
    if (!c->ssh_ready) return fgpu_fail(FGPU_ERR_STATE, "SSH matrix not set");
    FGPU_HIP(hipSetDevice(c->device));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    FGPU_HIP(hipMemcpy(x, c->ssh_x, sizeof(double) * c->ssh_n, hipMemcpyDeviceToHost));
    return FGPU_OK;
}
