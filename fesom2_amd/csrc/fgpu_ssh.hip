// SSH elliptic operator and preconditioned conjugate gradients for gfx950.
//
// * init_stiff_mat_ale (oce_ale.F90:1393-1582): one thread per matrix row
//   gathers the edge/element contributions of its row in ascending edge
//   order (the reference's serial order), so the assembly is deterministic
//   and free of atomics.
// * ssh_solve_preconditioner (solver.F90:31-117): one thread per row.
// * ssh_solve_cg (solver.F90:120-385): the iteration runs on the device as a
//   captured graph of small kernels; convergence, breakdown and the iteration
//   cap are decided on the device and later kernels of a replay return at once.
//   Iterates, residual and dot products are FP64.
#include <chrono>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "fgpu_internal.h"

#define SBLK 256

// Rows of a block. With a multiple of eight blocks, the blocks b with
// b % 8 == x (one XCD) take the x-th contiguous eighth of the rows, so that the
// neighbour gathers of nearby rows hit that XCD's L2; otherwise plain
// grid-stride rows. Every row in [0, n) is visited exactly once either way.
struct CGRows { int row, end, step; };
__device__ __forceinline__ CGRows cg_rows(int n, int nb)
{
    CGRows q;
    if (nb % NXCD == 0) {
        const int per = nb / NXCD, x = (int)(blockIdx.x % NXCD), j = (int)(blockIdx.x / NXCD);
        const int chunk = (n + NXCD - 1) / NXCD;
        q.row = x * chunk + j * SBLK + (int)threadIdx.x;
        q.end = min(n, (x + 1) * chunk);
        q.step = per * SBLK;
    } else {
        q.row = (int)(blockIdx.x * SBLK + threadIdx.x);
        q.end = n;
        q.step = nb * SBLK;
    }
    return q;
}

// ------------------------------------------------------------------------------
// assembly
// ------------------------------------------------------------------------------
// One CSR row per thread, its edges in ascending order and each edge's
// elements in order: init_stiff_mat_ale (oce_ale.F90:1393-1582) with the
// element depth zbar_e_bot - zbar_e_srf and the areasvol/dt diagonal; with dhe
// set, update_stiff_mat_ale (oce_ale.F90:1683-1792): -dhe(elem) in place of the
// depth, added to the current values.
__global__ void __launch_bounds__(SBLK) k_stiff_rows(DevMesh m, const int *rowptr, const int *colind, double *val,
                                                     double dt, double factor, const double *dhe)
{
    const int r = blockIdx.x * SBLK + threadIdx.x;
    if (r >= m.N2) return;
    const int p0 = rowptr[r], p1 = rowptr[r + 1];
    if (!dhe)
        for (int p = p0; p < p1; ++p) val[p] = 0.0;
    for (int pp = m.ne_ptr[r]; pp < m.ne_ptr[r + 1]; ++pp) {
        const int q = m.ne_idx[pp];
        const int ed = q >> 1, side = q & 1;
        for (int i = 0; i < 2; ++i) {
            const int el = m.edge_tri[2 * ed + i];
            if (el < 0) continue;
            const int *en = m.elem_nodes + 3 * (size_t)el;
            const double *gs = m.gsca + 6 * (size_t)el;
            const double hh = dhe ? -dhe[el] : m.zbar_e_bot[el] - m.zbar_e_srf[el];
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
    if (!dhe && m.ulevn[r] <= 1) val[p0] = val[p0] + m.areasvol[(size_t)r * m.NL + m.ulevn[r] - 1] / dt;
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
// CG as a hipGraph of small kernels (a grid-wide barrier costs far more than a
// kernel boundary on this GPU). One iteration of the reference loop is split
//   k_cg_a(i): convergence test of iteration i-1, beta, p_i = z + beta p_{i-1},
//              q = A p_i, partial p.q
//   k_cg_b(i): alpha, x and r update, z = M^-1 r, partial r.z and r.r
// Every block re-adds the per-block partials of the previous kernel in the
// same fixed order, so all blocks hold identical scalars and take identical
// branches. p is double-buffered (iteration parity) because k_cg_a reads
// p_{i-1} of neighbour rows while writing p_i; q = A p_i recomputes the
// neighbours' p_i from z and p_{i-1} with the same expression.
// ------------------------------------------------------------------------------
struct CGState {
    int iter;       // completed iterations (written by k_cg_b / init)
    int par;        // parity of the iteration in flight (written by k_cg_a)
    int done;       // 0 running, 1 converged, 2 breakdown, 3 maxiter
    int nbreak, nneg, pad;
    int itp[2];         // k_cg1: iteration number of the launch of each parity
    double s_oldp[4];   // k_cg1: gamma and alpha of the previous iteration, per parity
    double s_old[2];
    double rtol, rr_last;
    double soltol;      // stopping tolerance and iteration cap of the solve in flight, written by the
    int maxiter, pad2;  // init kernel: one captured iteration graph serves every (soltol, maxiter)
};

struct CGArgs {
    int n, variant, nb;
    const int *rowptr, *colind;
    const double *val, *pr, *dinv;
    double *x;
    const double *rhs;
    double *r, *z, *p0, *p1, *q;
    // single-reduction iteration: per row the record {r, w} {s_prev, D^-1}
    // (two double2), one array per iteration parity
    double2 *rec0, *rec1;
    double *part;   // 8 slots x nb: [0] b.b / p.q  [1] r.z  [2] r.r; single-reduction: [2p..2p+1], [4+p], [6]
    CGState *st;
};

#define SGB 8   // row entries whose column loads and gathers are in flight together

// sum over row r of v[p] * x[colind[p]] in entry order; the loads of SGB
// entries are issued before any is added
__device__ __forceinline__ double row_dot(const int *rowptr, const int *colind, const double *v, const double *x, int r)
{
    double s = 0.0;
    const int p0 = rowptr[r], p1 = rowptr[r + 1];
    for (int pb = p0; pb < p1; pb += SGB) {
        double xv[SGB], vv[SGB];
#pragma unroll
        for (int u = 0; u < SGB; ++u) {
            const int p = pb + u < p1 ? pb + u : p0;
            vv[u] = v[p];
            xv[u] = x[colind[p]];
        }
#pragma unroll
        for (int u = 0; u < SGB; ++u)
            if (pb + u < p1) s = s + vv[u] * xv[u];
    }
    return s;
}

// thread-strided sum of v[0..n): thread t adds v[t], v[t + SBLK], ... in that
// order, with four loads in flight
__device__ __forceinline__ double strided_sum(const double *v, int n)
{
    double a = 0.0;
    for (int base = threadIdx.x; base < n; base += 4 * SBLK) {
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = base + u * SBLK < n ? v[base + u * SBLK] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (base + u * SBLK < n) a += x[u];
    }
    return a;
}

// block reduction of up to three values: butterfly within each wave, then the
// four wave sums in wave order; the result (lane 0's sums) is valid in every
// thread and identical in every block for identical inputs
__device__ void block_sum3(double &a, double &b, double &c)
{
    __shared__ double sh[3][SBLK / 64];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        c += __shfl_xor(c, off);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[0][w] = a; sh[1][w] = b; sh[2][w] = c; }
    __syncthreads();
    a = sh[0][0]; b = sh[1][0]; c = sh[2][0];
    for (int i = 1; i < SBLK / 64; ++i) { a += sh[0][i]; b += sh[1][i]; c += sh[2][i]; }
    __syncthreads();
}

// every block adds the nb partials of slot s0 (and s0+1 if two) in one fixed order
__device__ void sum_partials(const CGArgs &A, int s0, bool two, double &a, double &b)
{
    double x = strided_sum(A.part + (size_t)s0 * A.nb, A.nb), z = 0.0;
    double y = two ? strided_sum(A.part + (size_t)(s0 + 1) * A.nb, A.nb) : 0.0;
    block_sum3(x, y, z);
    a = x;
    b = y;
}

__device__ __forceinline__ double apply_m(const CGArgs &A, int row)
{
    if (A.variant == 2) return A.dinv[row] * A.r[row];
    return row_dot(A.rowptr, A.colind, A.pr, A.r, row);
}

// r = b - A x; (Jacobi) z = D^-1 r; partials b.b and r.z; state reset
__global__ void __launch_bounds__(SBLK) k_cg_init(CGArgs A, double soltol, int maxiter)
{
    double bb = 0.0, rz = 0.0, dum = 0.0;
    for (CGRows R = cg_rows(A.n, A.nb); R.row < R.end; R.row += R.step) {
        const int row = R.row;
        const double rr = A.rhs[row] - row_dot(A.rowptr, A.colind, A.val, A.x, row);
        A.r[row] = rr;
        bb += A.rhs[row] * A.rhs[row];
        if (A.variant == 2) {
            const double zz = A.dinv[row] * rr;
            A.z[row] = zz;
            rz += rr * zz;
        }
    }
    block_sum3(bb, rz, dum);
    if (threadIdx.x == 0) {
        A.part[blockIdx.x] = bb;
        A.part[(size_t)A.nb + blockIdx.x] = rz;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        CGState s{};
        s.soltol = soltol;
        s.maxiter = maxiter;
        *A.st = s;
    }
}

// sparse M^-1 only: z = M^-1 r (needs r of neighbour rows), partial r.z
__global__ void __launch_bounds__(SBLK) k_cg_init_z(CGArgs A)
{
    double rz = 0.0, d1 = 0.0, d2 = 0.0;
    for (CGRows R = cg_rows(A.n, A.nb); R.row < R.end; R.row += R.step) {
        const int row = R.row;
        const double zz = apply_m(A, row);
        A.z[row] = zz;
        rz += A.r[row] * zz;
    }
    block_sum3(rz, d1, d2);
    if (threadIdx.x == 0) A.part[(size_t)A.nb + blockIdx.x] = rz;
}

__global__ void __launch_bounds__(SBLK) k_cg_a(CGArgs A)
{
    CGState *st = A.st;
    if (st->done) return;
    const int it = st->iter;          // completed iterations
    const int par = (it + 1) & 1;     // parity of iteration it+1
    double be = 0.0, s_new;
    const bool first = (it == 0);
    if (first) {
        double bb, rz;
        sum_partials(A, 0, true, bb, rz);
        const double rtol = st->soltol * sqrt(bb / (double)A.n);
        s_new = rz;
        if (blockIdx.x == 0 && threadIdx.x == 0) { st->rtol = rtol; st->rr_last = 0.0; }
    } else {
        double rz, rr;
        sum_partials(A, 1, true, rz, rr);
        const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
        if (lead) st->rr_last = rr;
        if (sqrt(rr / (double)A.n) < st->rtol) {
            if (lead) st->done = 1;
            return;
        }
        if (rz < 0.0 && lead) st->nneg += 1;
        const double s_prev = st->s_old[par ^ 1];
        if (s_prev == 0.0) {
            if (lead) { st->nbreak += 1; st->done = 2; }
            return;
        }
        be = rz / s_prev;
        s_new = rz;
    }
    if (it >= st->maxiter) {
        if (blockIdx.x == 0 && threadIdx.x == 0) st->done = 3;
        return;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) { st->s_old[par] = s_new; st->par = par; }
    const double *pold = par ? A.p0 : A.p1;
    double *pnew = par ? A.p1 : A.p0;
    double pq = 0.0, d1 = 0.0, d2 = 0.0;
    for (CGRows R = cg_rows(A.n, A.nb); R.row < R.end; R.row += R.step) {
        const int row = R.row;
        double s = 0.0;
        const int p0 = A.rowptr[row], p1 = A.rowptr[row + 1];
        for (int pb = p0; pb < p1; pb += SGB) {
            double pv[SGB], vv[SGB];
#pragma unroll
            for (int u = 0; u < SGB; ++u) {
                const int p = pb + u < p1 ? pb + u : p0;
                const int j = A.colind[p];
                vv[u] = A.val[p];
                pv[u] = first ? A.z[j] : A.z[j] + be * pold[j];
            }
#pragma unroll
            for (int u = 0; u < SGB; ++u)
                if (pb + u < p1) s = s + vv[u] * pv[u];
        }
        const double pr_ = first ? A.z[row] : A.z[row] + be * pold[row];
        pnew[row] = pr_;
        A.q[row] = s;
        pq += pr_ * s;
    }
    block_sum3(pq, d1, d2);
    if (threadIdx.x == 0) A.part[blockIdx.x] = pq;
}

__global__ void __launch_bounds__(SBLK) k_cg_b(CGArgs A, int fuse_z)
{
    CGState *st = A.st;
    if (st->done) return;
    const int par = st->par;
    double s_aux, dum;
    sum_partials(A, 0, false, s_aux, dum);
    if (s_aux <= 0.0 || s_aux != s_aux) {
        if (blockIdx.x == 0 && threadIdx.x == 0) { st->nbreak += 1; st->done = 2; st->iter += 1; }
        return;
    }
    const double al = st->s_old[par] / s_aux;
    const double *p = par ? A.p1 : A.p0;
    double rz = 0.0, rr = 0.0, d2 = 0.0;
    for (CGRows R = cg_rows(A.n, A.nb); R.row < R.end; R.row += R.step) {
        const int row = R.row;
        A.x[row] = A.x[row] + al * p[row];
        const double r_ = A.r[row] - al * A.q[row];
        A.r[row] = r_;
        if (fuse_z) {
            const double zz = A.dinv[row] * r_;
            A.z[row] = zz;
            rz += r_ * zz;
            rr += r_ * r_;
        }
    }
    if (fuse_z) {
        block_sum3(rz, rr, d2);
        if (threadIdx.x == 0) {
            A.part[(size_t)A.nb + blockIdx.x] = rz;
            A.part[(size_t)2 * A.nb + blockIdx.x] = rr;
        }
        if (blockIdx.x == 0 && threadIdx.x == 0) st->iter += 1;
    }
}

// sparse M^-1: z = M^-1 r after the r update, partials r.z and r.r
__global__ void __launch_bounds__(SBLK) k_cg_c(CGArgs A)
{
    CGState *st = A.st;
    if (st->done) return;
    double rz = 0.0, rr = 0.0, d2 = 0.0;
    for (CGRows R = cg_rows(A.n, A.nb); R.row < R.end; R.row += R.step) {
        const int row = R.row;
        const double zz = apply_m(A, row);
        A.z[row] = zz;
        rz += A.r[row] * zz;
        rr += A.r[row] * A.r[row];
    }
    block_sum3(rz, rr, d2);
    if (threadIdx.x == 0) {
        A.part[(size_t)A.nb + blockIdx.x] = rz;
        A.part[(size_t)2 * A.nb + blockIdx.x] = rr;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) st->iter += 1;
}

// ------------------------------------------------------------------------------
// Diagonally preconditioned CG with one reduction per iteration (Chronopoulos &
// Gear's arrangement of the same recurrences): with z = D^-1 r and w = A z,
//   gamma = r.z, delta = w.z, beta = gamma/gamma_prev,
//   alpha = gamma / (delta - beta gamma / alpha_prev),
//   s = w + beta s, p = z + beta p, x += alpha p, r -= alpha s.
// One kernel per iteration: it adds the previous iteration's partial sums,
// decides convergence on r.r like the reference, updates its rows and forms
// w = A z of the new residual, recomputing each neighbour's new z from the
// previous iteration's (double-buffered) r, w and s. Partial sums are double-
// buffered by iteration parity as well.
// ------------------------------------------------------------------------------
__device__ __forceinline__ double cg1_slot(const CGArgs &A, int slot)
{
    double x = strided_sum(A.part + (size_t)slot * A.nb, A.nb), d1 = 0.0, d2 = 0.0;
    block_sum3(x, d1, d2);
    return x;
}

// r0 = b - A x0 into the parity-0 records {r0, w0 (set next)} {s_-1 = 0, D^-1};
// partials b.b (slot 6) and r.r (slot 4); state reset
__global__ void __launch_bounds__(SBLK) k_cg1_init_r(CGArgs A, double soltol, int maxiter)
{
    double bb = 0.0, rr = 0.0, dum = 0.0;
    for (CGRows R = cg_rows(A.n, A.nb); R.row < R.end; R.row += R.step) {
        const int row = R.row;
        const double r = A.rhs[row] - row_dot(A.rowptr, A.colind, A.val, A.x, row);
        A.rec0[2 * row] = make_double2(r, 0.0);
        A.rec0[2 * row + 1] = make_double2(0.0, A.dinv[row]);
        bb += A.rhs[row] * A.rhs[row];
        rr += r * r;
    }
    block_sum3(bb, rr, dum);
    if (threadIdx.x == 0) {
        A.part[(size_t)6 * A.nb + blockIdx.x] = bb;
        A.part[(size_t)4 * A.nb + blockIdx.x] = rr;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        CGState s{};
        s.soltol = soltol;
        s.maxiter = maxiter;
        *A.st = s;
    }
}

// w0 = A z0 with z0 = D^-1 r0 into the parity-0 records; partials gamma0,
// delta0 (slots 0, 1); p_{-1} = 0
__global__ void __launch_bounds__(SBLK) k_cg1_init_w(CGArgs A)
{
    double g = 0.0, d = 0.0, dum = 0.0;
    double *rc = (double *)A.rec0;   // row i: rc[4i] = r, rc[4i+1] = w, rc[4i+3] = D^-1
    for (CGRows R = cg_rows(A.n, A.nb); R.row < R.end; R.row += R.step) {
        const int row = R.row;
        double w = 0.0;
        const int p0 = A.rowptr[row], p1 = A.rowptr[row + 1];
        for (int pb = p0; pb < p1; pb += SGB) {
            double zv[SGB], vv[SGB];
#pragma unroll
            for (int u = 0; u < SGB; ++u) {
                const int p = pb + u < p1 ? pb + u : p0;
                const int j = A.colind[p];
                vv[u] = A.val[p];
                zv[u] = rc[4 * (size_t)j + 3] * rc[4 * (size_t)j];
            }
#pragma unroll
            for (int u = 0; u < SGB; ++u)
                if (pb + u < p1) w = w + vv[u] * zv[u];
        }
        const double r = rc[4 * (size_t)row];
        const double z = rc[4 * (size_t)row + 3] * r;
        rc[4 * (size_t)row + 1] = w;   // only this row's w word: other blocks read r and D^-1
        A.p0[row] = 0.0;   // p_{-1}
        g += r * z;
        d += w * z;
    }
    block_sum3(g, d, dum);
    if (threadIdx.x == 0) {
        A.part[blockIdx.x] = g;
        A.part[(size_t)A.nb + blockIdx.x] = d;
    }
}

// iteration i reads the records of parity i & 1, {r_i, w_i} {s_{i-1}, D^-1},
// and writes {r_{i+1}, w_{i+1}} {s_i, D^-1} into the other parity's. par = i & 1
// comes with the launch (every graph replay runs an even number of iterations
// from an even one), and i and the previous scalars are read from the slots of
// that parity, so a block dispatched after others have finished still reads
// this iteration's values.
__global__ void __launch_bounds__(SBLK) k_cg1(CGArgs A, int par)
{
    CGState *st = A.st;
    if (st->done) return;
    const int it = st->itp[par];
    double gam = strided_sum(A.part + (size_t)(2 * par) * A.nb, A.nb);
    double del = strided_sum(A.part + (size_t)(2 * par + 1) * A.nb, A.nb);
    double rr = strided_sum(A.part + (size_t)(4 + par) * A.nb, A.nb);
    block_sum3(gam, del, rr);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    double rtol;
    if (it == 0) {
        rtol = st->soltol * sqrt(cg1_slot(A, 6) / (double)A.n);
        if (lead) st->rtol = rtol;
    } else {
        rtol = st->rtol;
    }
    if (lead) st->rr_last = rr;
    if (sqrt(rr / (double)A.n) < rtol) {
        if (lead) st->done = 1;
        return;
    }
    if (it >= st->maxiter) {
        if (lead) st->done = 3;
        return;
    }
    if (gam < 0.0 && lead) st->nneg += 1;
    double be = 0.0, den = del;
    if (it > 0) {
        const double gprev = st->s_oldp[2 * par], aprev = st->s_oldp[2 * par + 1];
        if (gprev == 0.0) {
            if (lead) { st->nbreak += 1; st->done = 2; }
            return;
        }
        be = gam / gprev;
        den = del - be * gam / aprev;
    }
    if (!(den > 0.0)) {
        if (lead) { st->nbreak += 1; st->done = 2; }
        return;
    }
    const double al = gam / den;
    const double2 *rc = par ? A.rec1 : A.rec0;
    double2 *rn = par ? A.rec0 : A.rec1;
    double g = 0.0, d = 0.0, r2 = 0.0;
    const CGRows R = cg_rows(A.n, A.nb);
    for (int row = R.row; row < R.end; row += R.step) {
        const double2 o0 = rc[2 * row], o1 = rc[2 * row + 1];   // {r, w}, {s_prev, D^-1}
        const double s_ = o0.y + be * o1.x;
        const double p_ = o1.y * o0.x + be * A.p0[row];
        A.p0[row] = p_;
        A.x[row] = A.x[row] + al * p_;
        const double r_ = o0.x - al * s_;
        // w = A z in entry order; the column, value and neighbour record loads
        // of SGB entries are issued before any is added. Each neighbour's z
        // comes from its record in two 16-byte loads.
        double w = 0.0;
        const int q0 = A.rowptr[row], q1 = A.rowptr[row + 1];
        for (int pb = q0; pb < q1; pb += SGB) {
            double zv[SGB], vv[SGB];
#pragma unroll
            for (int u = 0; u < SGB; ++u) {
                const int p = pb + u < q1 ? pb + u : q0;
                const int j = A.colind[p];
                vv[u] = A.val[p];
                const double2 a = rc[2 * j], b = rc[2 * j + 1];
                zv[u] = b.y * (a.x - al * (a.y + be * b.x));
            }
#pragma unroll
            for (int u = 0; u < SGB; ++u)
                if (pb + u < q1) w = w + vv[u] * zv[u];
        }
        rn[2 * row] = make_double2(r_, w);
        rn[2 * row + 1] = make_double2(s_, o1.y);
        const double z_ = o1.y * r_;
        g += r_ * z_;
        d += w * z_;
        r2 += r_ * r_;
    }
    block_sum3(g, d, r2);
    const int np = par ^ 1;
    if (threadIdx.x == 0) {
        A.part[(size_t)(2 * np) * A.nb + blockIdx.x] = g;
        A.part[(size_t)(2 * np + 1) * A.nb + blockIdx.x] = d;
        A.part[(size_t)(4 + np) * A.nb + blockIdx.x] = r2;
    }
    // the next iteration's number and scalars go to the other parity's slots,
    // which no block of this launch reads
    if (lead) {
        st->s_oldp[2 * np] = gam;
        st->s_oldp[2 * np + 1] = al;
        st->itp[np] = it + 1;
        st->iter = it + 1;
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
    double **vs[] = {&c->ssh_x, &c->ssh_x0, &c->ssh_rhs, &c->ssh_r, &c->ssh_z, &c->ssh_p, &c->ssh_p1, &c->ssh_q,
                     &c->ssh_diag_inv};
    for (double **v : vs)
        if (!*v && (rc = fgpu_alloc(c, (void **)v, n * sizeof(double)))) return rc;
    for (double **v : {&c->ssh_rec0, &c->ssh_rec1})
        if (!*v && (rc = fgpu_alloc(c, (void **)v, 4 * n * sizeof(double)))) return rc;
    int blocks = (int)((n + SBLK - 1) / SBLK);
    if (blocks > 512) blocks = 512;
    c->ssh_blocks = blocks;
    if (!c->ssh_part && (rc = fgpu_alloc(c, (void **)&c->ssh_part, sizeof(double) * 8 * (size_t)blocks))) return rc;
    if (!c->ssh_out && (rc = fgpu_alloc(c, (void **)&c->ssh_out, sizeof(CGState)))) return rc;
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
                       c->ssh_colind, c->ssh_val, dt, factor, (const double *)nullptr);
    FGPU_HIP(hipGetLastError());
    FGPU_HIP(hipStreamSynchronize(c->stream));
    c->ssh_ready = true;
    c->ssh_pr_ready = false;
    return FGPU_OK;
}

extern "C" int fgpu_update_stiff_mat_ale(fgpu_context *c, const double *dhe, double dt, double g, double alpha,
                                         double theta)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null argument");
    if (!dhe && !c->hbar_valid)
        return fgpu_fail(FGPU_ERR_ARG, "dhe is NULL and fgpu_compute_hbar_ale has not run");
    if (!c->mesh_ready) return fgpu_fail(FGPU_ERR_STATE, "mesh not initialised");
    if (!c->ssh_ready || c->ssh_n != c->N2) return fgpu_fail(FGPU_ERR_STATE, "SSH matrix not assembled (fgpu_init_stiff_mat_ale)");
    FGPU_HIP(hipSetDevice(c->device));
    if (!c->vv_dhe) {   // allocated on the first call, kept by the context
        int rc = fgpu_alloc(c, (void **)&c->vv_dhe, sizeof(double) * (size_t)c->F2);
        if (rc) return rc;
    }
    double *d_dhe = c->vv_dhe;   // NULL dhe: the device's, from fgpu_compute_hbar_ale
    if (dhe) FGPU_HIP(hipMemcpyAsync(d_dhe, dhe, sizeof(double) * (size_t)c->F2, hipMemcpyHostToDevice, c->stream));
    const double factor = g * dt * alpha * theta;
    hipLaunchKernelGGL(k_stiff_rows, dim3((c->N2 + SBLK - 1) / SBLK), dim3(SBLK), 0, c->stream, c->m, c->ssh_rowptr,
                       c->ssh_colind, c->ssh_val, dt, factor, (const double *)d_dhe);
    FGPU_HIP(hipGetLastError());
    FGPU_HIP(hipStreamSynchronize(c->stream));
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

// a preconditioner built by the caller (mesh%ssh_stiff%pr_values of
// ssh_solve_preconditioner) in place of the device's build; the diagonal of a
// variant-2 (Jacobi) M^-1 is taken from the first entry of each row
extern "C" int fgpu_ssh_set_preconditioner(fgpu_context *c, int variant, const double *pr_values)
{
    if (!c || !pr_values) return fgpu_fail(FGPU_ERR_ARG, "null argument");
    if (!c->ssh_ready) return fgpu_fail(FGPU_ERR_STATE, "SSH matrix not set");
    if (variant < 0 || variant > 2) return fgpu_fail(FGPU_ERR_ARG, "precond_variant must be 0, 1 or 2");
    FGPU_HIP(hipSetDevice(c->device));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    FGPU_HIP(hipMemcpy(c->ssh_pr, pr_values, sizeof(double) * c->ssh_nnz, hipMemcpyHostToDevice));
    std::vector<double> dinv(c->ssh_n);
    for (int r = 0; r < c->ssh_n; ++r) dinv[r] = pr_values[c->h_ssh_rowptr[r]];
    FGPU_HIP(hipMemcpy(c->ssh_diag_inv, dinv.data(), sizeof(double) * c->ssh_n, hipMemcpyHostToDevice));
    c->ssh_variant = variant;
    c->ssh_pr_ready = true;
    return FGPU_OK;
}

static CGArgs cg_args(fgpu_context *c)
{
    CGArgs A;
    A.n = c->ssh_n;
    A.variant = c->ssh_variant;
    A.nb = c->ssh_blocks;
    A.rowptr = c->ssh_rowptr;
    A.colind = c->ssh_colind;
    A.val = c->ssh_val;
    A.pr = c->ssh_pr;
    A.dinv = c->ssh_diag_inv;
    A.x = c->ssh_x;
    A.rhs = c->ssh_rhs;
    A.r = c->ssh_r;
    A.z = c->ssh_z;
    A.p0 = c->ssh_p;
    A.p1 = c->ssh_p1;
    A.q = c->ssh_q;
    A.rec0 = (double2 *)c->ssh_rec0;
    A.rec1 = (double2 *)c->ssh_rec1;
    A.part = c->ssh_part;
    A.st = (CGState *)c->ssh_out;
    return A;
}

#define CG_GRAPH_ITERS 8

static int enqueue_iters(fgpu_context *c, const CGArgs &A, int iters)
{
    const dim3 g(A.nb), b(SBLK);
    if (A.variant == 2 && c->ssh_single_reduction) {
        for (int i = 0; i < iters; ++i) hipLaunchKernelGGL(k_cg1, g, b, 0, c->stream, A, i & 1);
        FGPU_HIP(hipGetLastError());
        return FGPU_OK;
    }
    for (int i = 0; i < iters; ++i) {
        hipLaunchKernelGGL(k_cg_a, g, b, 0, c->stream, A);
        if (A.variant == 2) {
            hipLaunchKernelGGL(k_cg_b, g, b, 0, c->stream, A, 1);
        } else {
            hipLaunchKernelGGL(k_cg_b, g, b, 0, c->stream, A, 0);
            hipLaunchKernelGGL(k_cg_c, g, b, 0, c->stream, A);
        }
    }
    FGPU_HIP(hipGetLastError());
    return FGPU_OK;
}

// graph of CG_GRAPH_ITERS iterations for the current variant. The tolerance
// and the iteration cap are read from the device state, which the init kernel
// of each solve writes, so the graph is captured once per variant and never
// replaced between solves that differ only in (soltol, maxiter).
static int cg_graph(fgpu_context *c, const CGArgs &A)
{
    const int key = A.variant + (c->ssh_single_reduction ? 16 : 0);
    if (c->ssh_graph && c->ssh_graph_variant == key) return FGPU_OK;
    if (c->ssh_graph) { (void)hipGraphExecDestroy(c->ssh_graph); c->ssh_graph = nullptr; }
    hipGraph_t g;
    FGPU_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = enqueue_iters(c, A, CG_GRAPH_ITERS);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return fgpu_fail(FGPU_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&c->ssh_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fgpu_fail(FGPU_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    c->ssh_graph_variant = key;
    return FGPU_OK;
}

// One solve from the x currently in ssh_x. Replays the iteration graph in
// batches sized by the previous solve's iteration count and polls the device
// flag between batches; kernels after convergence return immediately.
static int launch_cg(fgpu_context *c, double soltol, int maxiter, int variant)
{
    CGArgs A = cg_args(c);
    A.variant = variant;
    int rc = cg_graph(c, A);
    if (rc) return rc;
    const dim3 g(A.nb), b(SBLK);
    if (A.variant == 2 && c->ssh_single_reduction) {
        hipLaunchKernelGGL(k_cg1_init_r, g, b, 0, c->stream, A, soltol, maxiter);
        hipLaunchKernelGGL(k_cg1_init_w, g, b, 0, c->stream, A);
    } else {
        hipLaunchKernelGGL(k_cg_init, g, b, 0, c->stream, A, soltol, maxiter);
        if (A.variant != 2) hipLaunchKernelGGL(k_cg_init_z, g, b, 0, c->stream, A);
    }
    FGPU_HIP(hipGetLastError());
    const int max_graphs = maxiter / CG_GRAPH_ITERS + 2;
    int expect = c->ssh_last_iters > 0 ? c->ssh_last_iters : 4 * CG_GRAPH_ITERS;
    int launched = 0;
    int batch = (expect + CG_GRAPH_ITERS - 1) / CG_GRAPH_ITERS;
    CGState hs;
    while (true) {
        for (int i = 0; i < batch && launched < max_graphs; ++i, ++launched)
            FGPU_HIP(hipGraphLaunch(c->ssh_graph, c->stream));
        FGPU_HIP(hipMemcpyAsync(&hs, c->ssh_out, sizeof(CGState), hipMemcpyDeviceToHost, c->stream));
        FGPU_HIP(hipStreamSynchronize(c->stream));
        if (hs.done || launched >= max_graphs) break;
        batch = 1;
    }
    if (!hs.done) return fgpu_fail(FGPU_ERR_STATE, "CG iteration graph did not terminate");
    c->ssh_last_iters = hs.iter;
    return FGPU_OK;
}

// plain CG from the current ssh_x with the given preconditioner variant (the
// fallback of the learned solve); summary of the device state afterwards
int fgpu_ssh_cg_from_x(fgpu_context *c, double soltol, int maxiter, int variant, FgpuCgSummary *out)
{
    int rc = launch_cg(c, soltol, maxiter, variant);
    if (rc) return rc;
    CGState hs;
    FGPU_HIP(hipMemcpy(&hs, c->ssh_out, sizeof(CGState), hipMemcpyDeviceToHost));
    out->iters = hs.iter;
    out->done = hs.done;
    out->nbreak = hs.nbreak;
    out->nneg = hs.nneg;
    out->rr_last = hs.rr_last;
    out->rtol = hs.rtol;
    return FGPU_OK;
}

static int read_stats(fgpu_context *c, fgpu_cg_stats *st, double ms)
{
    CGState hs;
    FGPU_HIP(hipMemcpy(&hs, c->ssh_out, sizeof(CGState), hipMemcpyDeviceToHost));
    if (st) {
        st->iters_last = hs.iter < 1 ? 1 : hs.iter;
        st->converged = hs.done == 1;
        st->nbreakdown = hs.nbreak;
        st->nnegrz = hs.nneg;
        st->resid_last = std::sqrt(hs.rr_last / (double)c->ssh_n);
        st->rtol_last = hs.rtol;
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
    rc = launch_cg(c, soltol, maxiter, c->ssh_variant);
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
        rc = launch_cg(c, soltol, maxiter, c->ssh_variant);
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

// ==============================================================================
// ssh_solve_cg without allocation: the work vectors (the two parities' records
// {r, w} {s, D^-1} and p of the single-reduction iteration, or r, z, p0, p1
// and q of the two-kernel one, sharing the area), the per-block partial sums and the CG state live in `scratch`, a static device
// array of FGPU_NOALLOC_MAX_ROWS rows that every GPU the library runs on holds
// until the process ends. x and rhs are the caller's device vectors. Kernels
// are launched directly (no graph instantiation), eight iterations between
// polls of the state, with the same kernels, launch order and iteration
// parities as fgpu_ssh_solve_cg, so both give the same x bit for bit.
// Calls that overlap on one GPU share `scratch` and race.
// ==============================================================================
#define NOALLOC_VECS 9   // doubles per row: two 4-double records and p
#define NOALLOC_PART (8 * 512)
static constexpr size_t NOALLOC_DOUBLES =
    (size_t)NOALLOC_VECS * FGPU_NOALLOC_MAX_ROWS + NOALLOC_PART + (sizeof(CGState) + 7) / 8;
static __device__ __align__(256) double scratch[NOALLOC_DOUBLES];

#define NOALLOC_MAX_DEVICES 64
static std::mutex g_scratch_mu;
static double *g_scratch_ptr[NOALLOC_MAX_DEVICES];   // device address of `scratch`, per device

// the device address of `scratch` on the current device (hipGetSymbolAddress
// once per device; the array's host-side name is not a device address)
static int scratch_on_device(int dev, double **out)
{
    if (dev < 0 || dev >= NOALLOC_MAX_DEVICES) return fgpu_fail(FGPU_ERR_ARG, "device index out of range");
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void *p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fgpu_fail(FGPU_ERR_HIP, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = (double *)p;
    }
    *out = g_scratch_ptr[dev];
    return FGPU_OK;
}

extern "C" int fgpu_ssh_solve_cg_noalloc(fgpu_context *c, double *x_dev, const double *rhs_dev, double soltol,
                                         int maxiter, fgpu_cg_stats *stats)
{
    int rc = cg_ready(c, soltol, maxiter);
    if (rc) return rc;
    if (!x_dev || !rhs_dev) return fgpu_fail(FGPU_ERR_ARG, "null vector");
    if (c->ssh_n > FGPU_NOALLOC_MAX_ROWS)
        return fgpu_fail(FGPU_ERR_ARG, "fgpu_ssh_solve_cg_noalloc: " + std::to_string(c->ssh_n) +
                                           " rows exceed the static work area of " +
                                           std::to_string(FGPU_NOALLOC_MAX_ROWS) + " rows");
    if (c->ssh_blocks * 8 > NOALLOC_PART) return fgpu_fail(FGPU_ERR_STATE, "partial-sum area too small");
    double *s = nullptr;
    if ((rc = scratch_on_device(c->device, &s))) return rc;
    const auto t0 = std::chrono::steady_clock::now();
    CGArgs A = cg_args(c);
    const size_t M = FGPU_NOALLOC_MAX_ROWS;
    A.x = x_dev;
    A.rhs = rhs_dev;
    A.rec0 = (double2 *)s;
    A.rec1 = (double2 *)(s + 4 * M);
    A.r = s;   // the two-kernel iteration's vectors reuse the records' area
    A.z = s + M;
    A.p1 = s + 2 * M;
    A.q = s + 3 * M;
    A.p0 = s + 8 * M;
    A.part = s + NOALLOC_VECS * M;
    A.st = (CGState *)(s + NOALLOC_VECS * M + NOALLOC_PART);
    const dim3 g(A.nb), b(SBLK);
    if (A.variant == 2 && c->ssh_single_reduction) {
        hipLaunchKernelGGL(k_cg1_init_r, g, b, 0, c->stream, A, soltol, maxiter);
        hipLaunchKernelGGL(k_cg1_init_w, g, b, 0, c->stream, A);
    } else {
        hipLaunchKernelGGL(k_cg_init, g, b, 0, c->stream, A, soltol, maxiter);
        if (A.variant != 2) hipLaunchKernelGGL(k_cg_init_z, g, b, 0, c->stream, A);
    }
    FGPU_HIP(hipGetLastError());
    const int max_batches = maxiter / CG_GRAPH_ITERS + 2;
    CGState hs;
    hs.done = 0;
    for (int batch = 0; batch < max_batches && !hs.done; ++batch) {
        if ((rc = enqueue_iters(c, A, CG_GRAPH_ITERS))) return rc;
        FGPU_HIP(hipMemcpyAsync(&hs, A.st, sizeof(CGState), hipMemcpyDeviceToHost, c->stream));
        FGPU_HIP(hipStreamSynchronize(c->stream));
    }
    if (!hs.done) return fgpu_fail(FGPU_ERR_STATE, "CG did not terminate");
    c->ssh_last_iters = hs.iter;
    const double ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    if (stats) {
        stats->iters_last = hs.iter < 1 ? 1 : hs.iter;
        stats->converged = hs.done == 1;
        stats->nbreakdown = hs.nbreak;
        stats->nnegrz = hs.nneg;
        stats->resid_last = std::sqrt(hs.rr_last / (double)c->ssh_n);
        stats->rtol_last = hs.rtol;
        stats->ms = ms;
    }
    return FGPU_OK;
}
