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

__device__ __forceinline__ double dmx(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmn(double a, double b) { return a < b ? a : b; }

struct VVArgs {
    double dt, maxcfl;
    int which_ale, use_wsplit;
    const double *uv, *hbar, *hbar_old, *water_flux;
    const double *zbar;                // zlevel: reference interface depths (nl)
    double min_hnode;                  // zlevel: thinnest layer, in units of its nominal thickness
    int lzstar;                        // zlevel: layers of the local zstar spread
    double *w, *w_e, *w_i, *cfl, *hnode_new;
    double *colmax;                    // max of CFL_z per column (nod2D)
    int *nan_columns;
};

// One block per node column, one lane per interface k = 1..NL.
//   1. thickness divergence of every layer, edges ascending, el(1) before el(2)
//      (oce_ale.F90:2175-2265)
//   2. cumulative sum from the bottom, division by the interface area
//      (oce_ale.F90:2277-2320)
//   3. zstar: the ssh change spread over the layers above nlevels_nod2D_min,
//      new layer thicknesses, surface water flux (oce_ale.F90:2539-2600);
//      zlevel: the change in the surface layer, or spread over the top
//      lzstar_lev layers by one lane (oce_ale.F90:2336-2420)
//   4. CFL_z (oce_ale.F90:2938-2966) and the W split (oce_ale.F90:3035-3047)
__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ double rld(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffll), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

#define VV_NB 4   // edges whose element loads are issued together

template <int LB>
__global__ void __launch_bounds__(LB) k_vert_vel(DevMesh m, VVArgs a)
{
    const int n = xcd_unit();
    if (n >= m.N2) return;
    __shared__ double sW[LB + 1], sH[LB + 1];
    __shared__ double sZl[33];   // zlevel: the new thicknesses of layers 1..lzstar_lev
    __shared__ int sBad;
    const int tid = threadIdx.x, k = tid + 1;
    if (tid == 0) sBad = 0;
    const int L = m.L, NL = m.NL;
    // the node's record, then one edge of its list per lane: slot, element
    // level ranges, elements and dual-face vectors
    const int4 nr = *(const int4 *)(m.node_rec + 4 * (size_t)n);
    const int p0 = nr.x, deg = nr.y, nzmin = nr.z, nzmax = nr.w - 1;
    const int j = tid & 63;
    int q = 0, el1 = 0, el2 = -1, nl1 = 0, nu1 = 1, nl2 = 0, nu2 = 1;
    double dx1 = 0.0, dy1 = 0.0, dx2 = 0.0, dy2 = 0.0;
    if (j < deg) {
        const int4 r0 = *(const int4 *)(m.ne_rec + 8 * (size_t)(p0 + j));
        const int4 r1 = *(const int4 *)(m.ne_rec + 8 * (size_t)(p0 + j) + 4);
        q = r0.x;
        nl1 = r1.x; nu1 = r1.y; nl2 = r1.z; nu2 = r1.w;
        const int e = q >> 1;
        const int2 et = *(const int2 *)(m.edge_tri + 2 * (size_t)e);
        el1 = et.x;
        el2 = et.y;
        const double2 c01 = *(const double2 *)(m.ecd + 4 * (size_t)e);
        const double2 c23 = *(const double2 *)(m.ecd + 4 * (size_t)e + 2);
        dx1 = c01.x; dy1 = c01.y; dx2 = c23.x; dy2 = c23.y;
    }
    // thickness divergence: edges in list (ascending) order, el(1) before
    // el(2); the element loads of VV_NB edges are in flight together
    double acc = 0.0;
    for (int j0 = 0; j0 < deg; j0 += VV_NB) {
        double2 u1[VV_NB], u2[VV_NB];
        double h1[VV_NB], h2[VV_NB];
#pragma unroll
        for (int b = 0; b < VV_NB; ++b) {
            const int jj = j0 + b < deg ? j0 + b : j0;
            const int e1 = rl(el1, jj), e2 = rl(el2, jj);
            const bool a1 = j0 + b < deg && k >= rl(nu1, jj) && k <= rl(nl1, jj);
            const bool a2 = j0 + b < deg && e2 >= 0 && k >= rl(nu2, jj) && k <= rl(nl2, jj);
            u1[b] = make_double2(0.0, 0.0);
            u2[b] = make_double2(0.0, 0.0);
            h1[b] = 0.0;
            h2[b] = 0.0;
            if (a1) {
                u1[b] = *(const double2 *)(a.uv + ((size_t)e1 * L + k - 1) * 2);
                h1[b] = m.helem[(size_t)e1 * L + k - 1];
            }
            if (a2) {
                u2[b] = *(const double2 *)(a.uv + ((size_t)e2 * L + k - 1) * 2);
                h2[b] = m.helem[(size_t)e2 * L + k - 1];
            }
        }
#pragma unroll
        for (int b = 0; b < VV_NB; ++b) {
            const int jj = j0 + b;
            if (jj >= deg) break;
            const int qq = rl(q, jj), e2 = rl(el2, jj);
            if (k >= rl(nu1, jj) && k <= rl(nl1, jj)) {
                const double c = (u1[b].y * rld(dx1, jj) - u1[b].x * rld(dy1, jj)) * h1[b];
                acc = (qq & 1) ? acc - c : acc + c;
            }
            if (e2 >= 0 && k >= rl(nu2, jj) && k <= rl(nl2, jj)) {
                const double c = -((u2[b].y * rld(dx2, jj) - u2[b].x * rld(dy2, jj)) * h2[b]);
                acc = (qq & 1) ? acc - c : acc + c;
            }
        }
    }
    if (tid < LB) sW[tid] = acc;
    if (tid == 0) sW[LB] = 0.0;
    __syncthreads();
    if (tid == 0) {   // serial, bottom up: the reference's association order
        double run = sW[nzmax];
        int nz = nzmax;
        for (; nz - 3 >= nzmin; nz -= 4) {   // four reads ahead of the dependent adds
            const double d0 = sW[nz - 1], d1 = sW[nz - 2], d2 = sW[nz - 3], d3 = sW[nz - 4];
            run = d0 + run; sW[nz - 1] = run;
            run = d1 + run; sW[nz - 2] = run;
            run = d2 + run; sW[nz - 3] = run;
            run = d3 + run; sW[nz - 4] = run;
        }
        for (; nz >= nzmin; --nz) { run = sW[nz - 1] + run; sW[nz - 1] = run; }
    }
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
    if (a.which_ale == FGPU_ALE_ZLEVEL && nzmin == 1) {   // uniform over the block
        const int lz = a.lzstar;
        sW[tid] = w;
        if (tid < lz) sZl[tid] = h;
        __syncthreads();
        if (tid == 0) {
            const double *ho = m.hnode + (size_t)n * L, *zb = a.zbar, *cprev = a.cfl + (size_t)n * NL;
            const double dt = a.dt, mh = a.min_hnode;
            const double dtot = a.hbar[n] - a.hbar_old[n];
            int zmax = m.nlevn_min[n] - 1;
            double mx[32], dist[32];
            if (dtot < 0.0 && ho[0] + dtot <= (zb[0] - zb[1]) * mh) {
                for (int j = 1; j <= lz; ++j) {
                    double v = (zb[j - 1] - zb[j]) * mh - ho[j - 1];
                    if (v >= 0.0) v = 0.0;
                    if (cprev[j - 1] >= 0.95) v = 0.0;   // the previous step's CFL_z
                    mx[j - 1] = v;
                }
                int nz = lz;   // first level whose pair sum is below the change (reference's cumsum_maxdhbar)
                for (int j = lz; j >= 1; --j) {
                    const double cj = j == 1 ? mx[0] : mx[j - 1] + mx[j - 2];
                    if (cj < dtot) nz = j;
                }
                double rest = dtot;
                zmax = min(nz, zmax - 1);
                for (int j = 1; j <= zmax; ++j) {
                    dist[j - 1] = dmx(rest, mx[j - 1]);
                    rest = rest - dist[j - 1];
                    rest = dmn(0.0, rest);
                }
                double dint = 0.0;
                for (int j = zmax; j >= 1; --j) {
                    dint = dint + dist[j - 1];
                    sW[j - 1] = sW[j - 1] - dint / dt;
                    sZl[j - 1] = ho[j - 1] + dist[j - 1];
                }
            } else {
                bool thin = false;
                for (int j = 2; j <= lz; ++j) thin = thin || ho[j - 1] != (zb[j - 1] - zb[j]);
                if (dtot > 0.0 && thin) {
                    int nz = 0;
                    for (int j = 1; j <= lz; ++j) {
                        mx[j - 1] = (zb[j - 1] - zb[j]) - ho[j - 1];
                        if (ho[j - 1] != (zb[j - 1] - zb[j])) nz = j;
                    }
                    mx[0] = 1000.0;
                    zmax = min(nz, zmax - 1);
                    double rest = dtot, dint = 0.0;
                    for (int j = zmax; j >= 1; --j) {
                        const double d = dmn(rest, mx[j - 1]);
                        rest = rest - d;
                        rest = dmx(0.0, rest);
                        dint = dint + d;
                        sW[j - 1] = sW[j - 1] - dint / dt;
                        sZl[j - 1] = ho[j - 1] + d;
                    }
                } else {
                    sW[0] = sW[0] - dtot / dt;
                    sZl[0] = ho[0] + dtot;
                }
            }
            sW[0] = sW[0] - a.water_flux[n];
        }
        __syncthreads();
        w = sW[tid];
        if (tid < lz) {
            h = sZl[tid];
            a.hnode_new[(size_t)n * L + tid] = h;
        }
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
    // column maximum of CFL_z (reduced over the columns by k_cfl_max)
    double cm = k <= NL ? cfl : 0.0;
    for (int off = 32; off > 0; off >>= 1) cm = fmax(cm, __shfl_xor(cm, off));
    __syncthreads();
    if ((tid & 63) == 0) sW[tid >> 6] = cm;
    __syncthreads();
    if (tid == 0) {
        for (int i = 1; i < LB / 64; ++i) cm = fmax(cm, sW[i]);
        a.colmax[n] = cm;
    }
}

// maximum over the column maxima: a block maximum per block, then one atomic
// max per block on the bit pattern (non-negative doubles and their bits order
// alike); the maximum does not depend on the order
__global__ void __launch_bounds__(256) k_cfl_max(const double *colmax, int n, unsigned long long *bits)
{
    __shared__ double sm[4];
    double cm = 0.0;
    for (int i = blockIdx.x * 256 + threadIdx.x; i < n; i += gridDim.x * 256) cm = fmax(cm, colmax[i]);
    for (int off = 32; off > 0; off >>= 1) cm = fmax(cm, __shfl_xor(cm, off));
    if ((threadIdx.x & 63) == 0) sm[threadIdx.x >> 6] = cm;
    __syncthreads();
    if (threadIdx.x == 0) {
        cm = fmax(fmax(sm[0], sm[1]), fmax(sm[2], sm[3]));
        atomicMax(bits, (unsigned long long)__double_as_longlong(cm));
    }
}

// Depth-integrated transport of alpha (u + u_rhs) through the two dual faces
// of edge e (oce_ale.F90:1844-1880): one block per edge, the layer terms of
// both elements computed by one lane per level, then each element's levels
// summed in order (lane 0: el(1), lane 1: el(2)) as the reference does.
template <int LB, bool HB>
__global__ void __launch_bounds__(LB) k_ssh_rhs_edge(DevMesh m, const double *uv, const double *uvr, double alpha,
                                                     double *ce)
{
    const int e = xcd_unit();
    if (e >= m.E2) return;
    __shared__ double s1[LB], s2[LB];
    const int tid = threadIdx.x, k = tid + 1, L = m.L;
    const int2 et = *(const int2 *)(m.edge_tri + 2 * (size_t)e);
    const int el1 = et.x, el2 = et.y;
    const int nu1 = m.ulev[el1], nl1 = m.nlev[el1] - 1;
    const int nu2 = el2 >= 0 ? m.ulev[el2] : 1, nl2 = el2 >= 0 ? m.nlev[el2] - 1 : 0;
    // HB: compute_hbar_ale's transport of u alone (oce_ale.F90:1988-2010)
    if (k >= nu1 && k <= nl1) {
        const double2 c = *(const double2 *)(m.ecd + 4 * (size_t)e);
        const size_t i = ((size_t)el1 * L + k - 1) * 2;
        const double2 u = *(const double2 *)(uv + i);
        if (HB) {
            s1[tid] = (u.y * c.x - u.x * c.y) * m.helem[(size_t)el1 * L + k - 1];
        } else {
            const double2 ur = *(const double2 *)(uvr + i);
            s1[tid] = alpha * ((u.y + ur.y) * c.x - (u.x + ur.x) * c.y) * m.helem[(size_t)el1 * L + k - 1];
        }
    }
    if (k >= nu2 && k <= nl2) {
        const double2 c = *(const double2 *)(m.ecd + 4 * (size_t)e + 2);
        const size_t i = ((size_t)el2 * L + k - 1) * 2;
        const double2 u = *(const double2 *)(uv + i);
        if (HB) {
            s2[tid] = (u.y * c.x - u.x * c.y) * m.helem[(size_t)el2 * L + k - 1];
        } else {
            const double2 ur = *(const double2 *)(uvr + i);
            s2[tid] = alpha * ((u.y + ur.y) * c.x - (u.x + ur.x) * c.y) * m.helem[(size_t)el2 * L + k - 1];
        }
    }
    __syncthreads();
    if (tid < 2) {
        const double *src = tid == 0 ? s1 : s2;
        const int lo = tid == 0 ? nu1 : nu2, hi = tid == 0 ? nl1 : nl2;
        // levels in order (the reference's association); the LDS reads of
        // eight levels are issued before their adds
        double c = 0.0;
        int nz = lo;
        for (; nz + 7 <= hi; nz += 8) {
            double v[8];
#pragma unroll
            for (int u = 0; u < 8; ++u) v[u] = src[nz - 1 + u];
#pragma unroll
            for (int u = 0; u < 8; ++u) c = tid == 0 ? c + v[u] : c - v[u];
        }
        for (; nz <= hi; ++nz) c = tid == 0 ? c + src[nz - 1] : c - src[nz - 1];
        const double c2 = __shfl(c, 1);
        if (tid == 0) ce[e] = c + c2;
    }
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

// compute_hbar_ale (oce_ale.F90:2013-2083): node gather of the edge
// transports in edge order, the freshwater flux off linfs, hbar_old = hbar and
// the new hbar at open-ocean nodes
__global__ void k_hbar_node(DevMesh m, const double *ce, const double *wf, int which_ale, double dt, double *old,
                            double *hbar, double *hbar_old)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.N2) return;
    double s = 0.0;
    for (int p = m.ne_ptr[n]; p < m.ne_ptr[n + 1]; ++p) {
        const int q = m.ne_idx[p];
        s = (q & 1) ? s - ce[q >> 1] : s + ce[q >> 1];
    }
    const int nzmin = m.ulevn[n];
    const double av = m.areasvol[(size_t)n * m.NL + nzmin - 1];
    if (which_ale != FGPU_ALE_LINFS && nzmin <= 1) s = s - wf[n] * av;
    old[n] = s;
    const double h0 = hbar[n];
    hbar_old[n] = h0;
    if (nzmin <= 1) hbar[n] = h0 + s * dt / av;   // cavity nodes keep hbar
}

// element thickness change dhe (oce_ale.F90:2084-2090)
__global__ void k_hbar_dhe(DevMesh m, const double *hbar, const double *hbar_old, double *dhe)
{
    const int el = blockIdx.x * blockDim.x + threadIdx.x;
    if (el >= m.F2) return;
    if (m.ulev[el] > 1) {
        dhe[el] = 0.0;
        return;
    }
    const int a = m.elem_nodes[3 * el], b = m.elem_nodes[3 * el + 1], c = m.elem_nodes[3 * el + 2];
    dhe[el] = ((hbar[a] - hbar_old[a]) + (hbar[b] - hbar_old[b]) + (hbar[c] - hbar_old[c])) / 3.0;
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
    if (which_ale != FGPU_ALE_LINFS && which_ale != FGPU_ALE_ZLEVEL && which_ale != FGPU_ALE_ZSTAR)
        return fgpu_fail(FGPU_ERR_ARG, "unknown which_ALE");
    if (which_ale == FGPU_ALE_ZLEVEL && !c->ale_zbar)
        return fgpu_fail(FGPU_ERR_STATE, "zlevel needs zbar, min_hnode and lzstar_lev (fgpu_set_ale_config)");
    if (!(dt > 0.0)) return fgpu_fail(FGPU_ERR_ARG, "dt must be > 0");
    if (c->nl > 128) return fgpu_fail(FGPU_ERR_ARG, "at most 128 interfaces per column");
    const size_t N2 = c->N2, NL = c->nl;
    VVArgs a{};
    a.dt = dt;
    a.maxcfl = wsplit_maxcfl;
    a.which_ale = which_ale;
    a.use_wsplit = use_wsplit ? 1 : 0;
    a.uv = c->d.uv;
    if (which_ale != FGPU_ALE_LINFS) {   // zstar and zlevel
        if (!c->forcing_ready) return fgpu_fail(FGPU_ERR_STATE, "zstar/zlevel need the surface water flux (fgpu_set_surface_forcing)");
        if (!hbar && !hbar_old) {   // the device's hbar, hbar_old from fgpu_compute_hbar_ale
            if (!c->hbar_valid) return fgpu_fail(FGPU_ERR_STATE, "zstar/zlevel need hbar and hbar_old (or fgpu_compute_hbar_ale first)");
        } else {
            if (!hbar || !hbar_old) return fgpu_fail(FGPU_ERR_ARG, "zstar/zlevel need both hbar and hbar_old");
            if ((rc = upload_to(c, &c->vv_hbar, hbar, N2))) return rc;
            if ((rc = upload_to(c, &c->vv_hbar_old, hbar_old, N2))) return rc;
        }
        a.hbar = c->vv_hbar;
        a.hbar_old = c->vv_hbar_old;
        a.water_flux = c->d.water_flux;
        a.zbar = c->ale_zbar;
        a.min_hnode = c->ale_min_hnode;
        a.lzstar = c->ale_lzstar;
    }
    if ((rc = ensure_const_buf(c, &c->d.w, N2 * NL))) return rc;
    if ((rc = ensure_const_buf(c, &c->d.w_e, N2 * NL))) return rc;
    if ((rc = ensure_const_buf(c, &c->d.w_i, N2 * NL))) return rc;
    if (!c->vv_cfl) {   // CFL_z starts at zero; zlevel reads the previous step's
        if ((rc = ensure_buf(c, &c->vv_cfl, N2 * NL))) return rc;
        FGPU_HIP(hipMemsetAsync(c->vv_cfl, 0, sizeof(double) * N2 * NL, c->stream));
    }
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
    if ((rc = ensure_buf(c, &c->vv_colmax, N2))) return rc;
    a.colmax = c->vv_colmax;
    a.nan_columns = (int *)((char *)c->vv_flags + 8);
    FGPU_HIP(hipMemsetAsync(c->vv_flags, 0, 16, c->stream));
    {
        ProfScope ps(c, "k_vert_vel");
        if (NL <= 64) hipLaunchKernelGGL(k_vert_vel<64>, dim3(xcd_grid(c->N2)), dim3(64), 0, c->stream, c->m, a);
        else hipLaunchKernelGGL(k_vert_vel<128>, dim3(xcd_grid(c->N2)), dim3(128), 0, c->stream, c->m, a);
        FGPU_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_cfl_max, dim3(256), dim3(256), 0, c->stream, (const double *)c->vv_colmax, c->N2,
                       (unsigned long long *)c->vv_flags);
    FGPU_HIP(hipGetLastError());
    if (which_ale != FGPU_ALE_LINFS && (rc = fgpu_compute_zn(c))) return rc;   // Z levels of the implicit diffusion
    unsigned long long flags[2] = {0ull, 0ull};
    FGPU_HIP(hipMemcpyAsync(flags, c->vv_flags, 16, hipMemcpyDeviceToHost, c->stream));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    if (c->prof && (rc = fgpu_prof_flush(c))) return rc;
    if (cflmax) *cflmax = __builtin_bit_cast(double, flags[0]);
    if (nan_columns) *nan_columns = (int)(flags[1] & 0xffffffffull);
    if (c->d.wsplit != (use_wsplit ? 1 : 0)) c->graph_valid = false;
    c->d.wsplit = use_wsplit ? 1 : 0;
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
    if (!uv_rhs) return fgpu_fail(FGPU_ERR_ARG, "null array");
    if (!ssh_rhs_old && !c->hbar_valid)
        return fgpu_fail(FGPU_ERR_STATE, "ssh_rhs_old is NULL and fgpu_compute_hbar_ale has not run");
    if (which_ale != FGPU_ALE_LINFS && which_ale != FGPU_ALE_ZLEVEL && which_ale != FGPU_ALE_ZSTAR)
        return fgpu_fail(FGPU_ERR_ARG, "unknown which_ALE");
    if (which_ale != FGPU_ALE_LINFS && !c->forcing_ready)
        return fgpu_fail(FGPU_ERR_STATE, "the full free surface needs the water flux (fgpu_set_surface_forcing)");
    const size_t N2 = c->N2, E2 = c->E2, F2 = c->F2, L = c->nl - 1;
    if ((rc = upload_to(c, &c->vv_uvr, uv_rhs, F2 * L * 2))) return rc;
    if (ssh_rhs_old && (rc = upload_to(c, &c->vv_old, ssh_rhs_old, N2))) return rc;   // else the device's
    if ((rc = ensure_buf(c, &c->vv_ce, E2))) return rc;
    if ((rc = ensure_buf(c, &c->vv_rhs, N2))) return rc;
    const double *wf = c->d.water_flux ? c->d.water_flux : c->vv_old;   // unread under linfs
    {
        ProfScope ps(c, "k_ssh_rhs_edge");
        if (c->nl - 1 <= 64)
            hipLaunchKernelGGL((k_ssh_rhs_edge<64, false>), dim3(xcd_grid(c->E2)), dim3(64), 0, c->stream, c->m, c->d.uv, c->vv_uvr,
                               alpha, c->vv_ce);
        else
            hipLaunchKernelGGL((k_ssh_rhs_edge<128, false>), dim3(xcd_grid(c->E2)), dim3(128), 0, c->stream, c->m, c->d.uv, c->vv_uvr,
                               alpha, c->vv_ce);
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

extern "C" int fgpu_compute_hbar_ale(fgpu_context *c, double dt, int which_ale, const double *hbar, double *hbar_out,
                                     double *hbar_old_out, double *ssh_rhs_old_out, double *dhe_out)
{
    int rc = check_dyn(c);
    if (rc) return rc;
    if (which_ale != FGPU_ALE_LINFS && which_ale != FGPU_ALE_ZLEVEL && which_ale != FGPU_ALE_ZSTAR)
        return fgpu_fail(FGPU_ERR_ARG, "unknown which_ALE");
    if (which_ale != FGPU_ALE_LINFS && !c->forcing_ready)
        return fgpu_fail(FGPU_ERR_STATE, "the full free surface needs the water flux (fgpu_set_surface_forcing)");
    if (!hbar && !c->hbar_valid) return fgpu_fail(FGPU_ERR_ARG, "hbar is NULL and no earlier hbar is on the device");
    const size_t N2 = c->N2, E2 = c->E2, F2 = c->F2;
    if (hbar && (rc = upload_to(c, &c->vv_hbar, hbar, N2))) return rc;
    if ((rc = ensure_buf(c, &c->vv_hbar_old, N2))) return rc;
    if ((rc = ensure_buf(c, &c->vv_old, N2))) return rc;
    if ((rc = ensure_buf(c, &c->vv_ce, E2))) return rc;
    if ((rc = ensure_buf(c, &c->vv_dhe, F2))) return rc;
    const double *wf = c->d.water_flux ? c->d.water_flux : c->vv_old;   // unread under linfs
    {
        ProfScope ps(c, "k_hbar_edge");
        if (c->nl - 1 <= 64)
            hipLaunchKernelGGL((k_ssh_rhs_edge<64, true>), dim3(xcd_grid(c->E2)), dim3(64), 0, c->stream, c->m, c->d.uv, c->d.uv,
                               1.0, c->vv_ce);
        else
            hipLaunchKernelGGL((k_ssh_rhs_edge<128, true>), dim3(xcd_grid(c->E2)), dim3(128), 0, c->stream, c->m, c->d.uv, c->d.uv,
                               1.0, c->vv_ce);
        FGPU_HIP(hipGetLastError());
    }
    hipLaunchKernelGGL(k_hbar_node, dim3((c->N2 + 255) / 256), dim3(256), 0, c->stream, c->m, (const double *)c->vv_ce,
                       wf, which_ale, dt, c->vv_old, c->vv_hbar, c->vv_hbar_old);
    FGPU_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_hbar_dhe, dim3((c->F2 + 255) / 256), dim3(256), 0, c->stream, c->m, (const double *)c->vv_hbar,
                       (const double *)c->vv_hbar_old, c->vv_dhe);
    FGPU_HIP(hipGetLastError());
    struct { double *dst; const double *src; size_t n; } cp[] = {
        {hbar_out, c->vv_hbar, N2}, {hbar_old_out, c->vv_hbar_old, N2}, {ssh_rhs_old_out, c->vv_old, N2},
        {dhe_out, c->vv_dhe, F2}};
    for (auto &x : cp)
        if (x.dst) FGPU_HIP(hipMemcpyAsync(x.dst, x.src, sizeof(double) * x.n, hipMemcpyDeviceToHost, c->stream));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    if (c->prof && (rc = fgpu_prof_flush(c))) return rc;
    c->hbar_valid = true;
    return FGPU_OK;
}

extern "C" int fgpu_set_ale_config(fgpu_context *c, const double *zbar, double min_hnode, int lzstar_lev)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    if (!zbar) return fgpu_fail(FGPU_ERR_ARG, "null zbar");
    if (lzstar_lev < 1 || lzstar_lev > 32 || lzstar_lev > c->nl - 1)
        return fgpu_fail(FGPU_ERR_ARG, "lzstar_lev must be in 1..min(32, nl-1)");
    if (!(min_hnode >= 0.0)) return fgpu_fail(FGPU_ERR_ARG, "min_hnode must be >= 0");
    FGPU_HIP(hipSetDevice(c->device));
    int rc = upload_to(c, &c->ale_zbar, zbar, (size_t)c->nl);
    if (rc) return rc;
    FGPU_HIP(hipStreamSynchronize(c->stream));
    c->ale_min_hnode = min_hnode;
    c->ale_lzstar = lzstar_lev;
    return FGPU_OK;
}
