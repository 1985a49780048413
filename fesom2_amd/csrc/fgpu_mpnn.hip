// Learned SSH preconditioner for gfx950: a message-passing network on the SSH
// matrix graph, used inside flexible CG, with its own training kernels.
//
// Network (width C = 64, three message-passing layers), for a residual r:
//   D = diag(A), r~ = D^-1/2 r, s = rms(r~), u = r~ / s
//   h0 = relu(u W0[0] + W0[1])                         FP32, stored bf16
//   m_l = Ahat h_{l-1}                                  FP32 sums, stored bf16
//   h_l = relu([h_{l-1}, m_l] W_l + b_l)               bf16 x bf16 -> FP32 on MFMA
//   y = h3 . w_out + b_out,  z = D^-1/2 (r~ + s y)      FP32 / FP64
// with Ahat_ij = a_ij / sqrt(a_ii a_jj) (off-diagonal entries of A). With the
// output weights at zero the network is the diagonal preconditioner.
//
// Layer kernels work on 64-node tiles (256 threads): the block gathers the
// neighbour features of its tile into LDS, then four waves run
// v_mfma_f32_32x32x16_bf16 over K = 128 (32 nodes x 32 channels per wave).
// Training is loss = mean (z - x)^T A (z - x) / (x^T A x) over right-hand
// sides b with solutions x; backward kernels produce per-block partial
// gradients (fixed block count) that one kernel adds up in block order, so
// gradients are deterministic. Adam runs on FP32 master weights.
//
// Flexible CG (FCG(1), Notay): p_k = z_k - (z_k.A p_{k-1}) / (p_{k-1}.A p_{k-1}) p_{k-1};
// vectors, residual and dot products in FP64. A solve that hits its
// iteration cap or breaks down continues with plain Jacobi CG from its x.
#include <cmath>
#include <random>

#include "fgpu_internal.h"

#define NNC 64
#define NNT 256
#define NNTILE 64
#define NNB 256   // blocks of the gradient kernels = partial rows

// parameter vector (FP32): W0(2,64) | W1(128,64) b1(64) | W2 b2 | W3 b3 | w_out(64) b_out(1)
#define OFF_W0 0
#define OFF_W(l) (128 + ((l) - 1) * (128 * 64 + 64))
#define OFF_B(l) (OFF_W(l) + 128 * 64)
#define OFF_WOUT (128 + 3 * (128 * 64 + 64))
#define OFF_BOUT (OFF_WOUT + 64)
#define NNP (OFF_BOUT + 1)

typedef unsigned short bf16_t;
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));

// partial-sum slots (double, nbd blocks each)
enum { SL_BB = 0, SL_RR, SL_RTRT, SL_ZR, SL_ZQ, SL_PQ, SL_XAX, SL_EAE, SL_JR, SL_JQ, SL_PQ1, NSLOT };

struct FCGState {
    int iter, done, nbreak, nneg;   // nneg: iterations that fell back to the diagonal direction
    double rtol, rr_last, s, zr;
    double soltol;                  // the solve's tolerance and iteration cap, written by k_fcg_init:
    int maxiter, pad;               // one captured iteration graph serves every (soltol, maxiter)
};

struct NNDev {
    int n = 0, nnz = 0, ntiles = 0, nbd = 0;
    float *w = nullptr, *wt = nullptr;       // Ahat per CSR entry and its transpose (0 on the diagonal)
    double *isq = nullptr;                   // 1/sqrt(a_ii)
    float *theta = nullptr, *gacc = nullptr, *am = nullptr, *av = nullptr;
    bf16_t *Wf = nullptr, *Wb = nullptr;     // 3 layers: [out 64][in 128] and [in 128][out 64]
    bf16_t *H = nullptr, *M = nullptr;       // 4 and 3 feature maps of n x 64
    float *u = nullptr, *y = nullptr;
    float *dH = nullptr, *dXs = nullptr, *dXm = nullptr;
    float *part = nullptr;                   // NNB x NNP
    double *dpart = nullptr;                 // NSLOT x nbd
    double *tpart = nullptr;                 // 4 x ntiles
    double *ae = nullptr, *zs = nullptr;     // A e and z of a training sample
    double *loss = nullptr;                  // loss accumulator
    FCGState *st = nullptr;
    double *samp_b = nullptr, *samp_x = nullptr;
    int nsamp_cap = 0;
    int adam_t = 0;
    bool ready = false, weights = false;
    hipGraphExec_t fcg_graph = nullptr;
    int fcg_last_iters = 0;
};

struct NNArgs {
    int n, ntiles, nbd;
    const int *rowptr, *colind;
    const double *val;
    const float *w, *wt;
    const double *isq;
    const float *theta;
    const bf16_t *Wf, *Wb;
    bf16_t *H, *M;
    float *u, *y, *dH, *dXs, *dXm, *part;
    double *dpart, *ae, *loss;
    double *tpart;  // 4 x ntiles per-tile partials of the last layer (FCG)
    FCGState *st;   // null outside FCG
};

__device__ __forceinline__ bf16_t f2bf(float f)
{
    unsigned int u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return (bf16_t)((u >> 16) | 0x40u);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (bf16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(bf16_t b) { return __uint_as_float(((unsigned int)b) << 16); }

__device__ __forceinline__ bool nn_done(const NNArgs &A) { return A.st && A.st->done; }

// block sum of three doubles (butterfly per wave, then the wave sums in order);
// valid in every thread, identical in every block for identical inputs
__device__ void nn_bsum3(double &a, double &b, double &c)
{
    __shared__ double sh[3][NNT / 64];
    for (int off = 32; off > 0; off >>= 1) {
        a += __shfl_xor(a, off);
        b += __shfl_xor(b, off);
        c += __shfl_xor(c, off);
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sh[0][w] = a; sh[1][w] = b; sh[2][w] = c; }
    __syncthreads();
    a = sh[0][0]; b = sh[1][0]; c = sh[2][0];
    for (int i = 1; i < NNT / 64; ++i) { a += sh[0][i]; b += sh[1][i]; c += sh[2][i]; }
    __syncthreads();
}

// thread-strided sum of v[0..n): thread t adds v[t], v[t + NNT], ... in that
// order, with four loads in flight
__device__ __forceinline__ double nn_strided_sum(const double *v, int n)
{
    double a = 0.0;
    for (int base = threadIdx.x; base < n; base += 4 * NNT) {
        double x[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) x[u] = base + u * NNT < n ? v[base + u * NNT] : 0.0;
#pragma unroll
        for (int u = 0; u < 4; ++u)
            if (base + u * NNT < n) a += x[u];
    }
    return a;
}

// three slots at once
__device__ void nn_slots3(const double *dpart, int nbd, int s0, int s1, int s2, double &a, double &b, double &c)
{
    a = nn_strided_sum(dpart + (size_t)s0 * nbd, nbd);
    b = s1 >= 0 ? nn_strided_sum(dpart + (size_t)s1 * nbd, nbd) : 0.0;
    c = s2 >= 0 ? nn_strided_sum(dpart + (size_t)s2 * nbd, nbd) : 0.0;
    nn_bsum3(a, b, c);
}

// the nbd partials of a slot, added in block order by every block
__device__ double nn_slot(const NNArgs &A, int slot)
{
    double x = nn_strided_sum(A.dpart + (size_t)slot * A.nbd, A.nbd), d1 = 0.0, d2 = 0.0;
    nn_bsum3(x, d1, d2);
    return x;
}

#define NSG 8   // matrix row entries whose loads are in flight together

__device__ __forceinline__ void nn_put3(const NNArgs &A, int s0, double a, int s1, double b, int s2, double c)
{
    nn_bsum3(a, b, c);
    if (threadIdx.x == 0) {
        if (s0 >= 0) A.dpart[(size_t)s0 * A.nbd + blockIdx.x] = a;
        if (s1 >= 0) A.dpart[(size_t)s1 * A.nbd + blockIdx.x] = b;
        if (s2 >= 0) A.dpart[(size_t)s2 * A.nbd + blockIdx.x] = c;
    }
}

// ------------------------------------------------------------------------------
// forward
// ------------------------------------------------------------------------------
// partial sums of (D^-1/2 r)^2 for the scale of a training sample
__global__ void __launch_bounds__(NNT) k_nn_rms(NNArgs A, const double *rin)
{
    double s = 0.0;
    for (int i = blockIdx.x * NNT + threadIdx.x; i < A.n; i += A.nbd * NNT) {
        const double rt = A.isq[i] * rin[i];
        s += rt * rt;
    }
    nn_put3(A, SL_RTRT, s, -1, 0.0, -1, 0.0);
}

#define LDX 136   // LDS row stride (bf16) of the 128-wide tiles
#define LDT 72    // LDS row stride (bf16) of the 64-wide tiles
#define NGB 8     // neighbours per gather batch

// h0 of a node: relu(u W0[0] + W0[1]) on channels c0..c0+7, packed bf16
__device__ __forceinline__ u32x4 nn_h0_pack(float u, const float *W0, int c0)
{
    unsigned int pk[4];
    for (int c = 0; c < 8; c += 2) {
        float p0 = u * W0[c0 + c] + W0[NNC + c0 + c];
        float p1 = u * W0[c0 + c + 1] + W0[NNC + c0 + c + 1];
        p0 = p0 > 0.f ? p0 : 0.f;
        p1 = p1 > 0.f ? p1 : 0.f;
        pk[c >> 1] = (unsigned int)f2bf(p0) | ((unsigned int)f2bf(p1) << 16);
    }
    const u32x4 v = {pk[0], pk[1], pk[2], pk[3]};
    return v;
}

__device__ __forceinline__ void nn_acc8(float *acc, float wv, const u32x4 a)
{
    const unsigned int wd[4] = {a.x, a.y, a.z, a.w};
    for (int c = 0; c < 4; ++c) {
        acc[2 * c] = acc[2 * c] + wv * bf2f((bf16_t)(wd[c] & 0xffffu));
        acc[2 * c + 1] = acc[2 * c + 1] + wv * bf2f((bf16_t)(wd[c] >> 16));
    }
}

// input layer over all nodes (grid-stride, nbd blocks): u and H[0]; in FCG mode
// first the convergence test on r.r, and the scale s for the later layers
__global__ void __launch_bounds__(NNT) k_nn_in(NNArgs A, const double *rin, double soltol)
{
    if (nn_done(A)) return;
    double rtrt, bb, rr;
    nn_slots3(A.dpart, A.nbd, SL_RTRT, A.st ? SL_BB : -1, A.st ? SL_RR : -1, rtrt, bb, rr);
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (A.st) {
        const double rtol = A.st->soltol * sqrt(bb / (double)A.n);
        if (lead) { A.st->rtol = rtol; A.st->rr_last = rr; }
        if (sqrt(rr / (double)A.n) < rtol) {
            if (lead) A.st->done = 1;
            return;
        }
    }
    double s = sqrt(rtrt / (double)A.n);
    if (!(s > 0.0)) s = 1.0;
    if (A.st && lead) A.st->s = s;
    // the input layer's weights in LDS: read once per block, not once per
    // channel pair behind every store of the loop below
    __shared__ float W0[2 * NNC];
    if (threadIdx.x < 2 * NNC) W0[threadIdx.x] = A.theta[OFF_W0 + threadIdx.x];
    __syncthreads();
    for (int i = blockIdx.x * NNT + threadIdx.x; i < A.n; i += A.nbd * NNT) {
        const float u = (float)(A.isq[i] * rin[i] / s);
        A.u[i] = u;
        for (int c8 = 0; c8 < NNC; c8 += 8) *(u32x4 *)(A.H + (size_t)i * NNC + c8) = nn_h0_pack(u, W0, c8);
    }
}

// Message-passing layer l (1..3), one 64-node tile per block (grid = ntiles):
//   H[l] = relu([H[l-1], Ahat H[l-1]] W_l + b_l), M[l-1] = Ahat H[l-1] (bf16)
// l = 1 also forms the input layer (H[0], u) from rin, and in FCG mode first
// decides convergence on r.r; l = 3 also forms y = h3 . w_out + b_out,
// z = D^-1/2 (r~ + s y) and the partial sums z.r, z.q, r.D^-1 r, (D^-1 r).q.
template <int LL>
__global__ void __launch_bounds__(NNT) k_nn_layer(NNArgs A, const double *rin, double *zout, const double *q,
                                                  double soltol)
{
    if (nn_done(A)) return;
    __shared__ __attribute__((aligned(16))) bf16_t sX[NNTILE * LDX];
    __shared__ __attribute__((aligned(16))) bf16_t sW[NNC * LDX];
    __shared__ float sY[2][NNTILE];
    bf16_t *sO = sX;   // output tile, written after the MFMA reads of sX
    const int t = threadIdx.x;
    double s = 1.0;
    if (LL == 3) {
        if (A.st) {
            s = A.st->s;
        } else {
            double rtrt, u1, u2;
            nn_slots3(A.dpart, A.nbd, SL_RTRT, -1, -1, rtrt, u1, u2);
            s = sqrt(rtrt / (double)A.n);
            if (!(s > 0.0)) s = 1.0;
        }
    }
    const bf16_t *Hin = A.H + (size_t)(LL - 1) * A.n * NNC;
    bf16_t *Hout = A.H + (size_t)LL * A.n * NNC;
    bf16_t *Mout = A.M + (size_t)(LL - 1) * A.n * NNC;
    const float *W0 = A.theta + OFF_W0;
    const bf16_t *Wf = A.Wf + (size_t)(LL - 1) * NNC * 128;
    for (int qq = t; qq < NNC * 16; qq += NNT) {   // W_l^T: 64 rows of 128
        const int row = qq >> 4, c8 = (qq & 15) * 8;
        *(u32x4 *)&sW[row * LDX + c8] = *(const u32x4 *)(Wf + row * 128 + c8);
    }
    const int wv = t >> 6, lane = t & 63;
    const int rt = wv >> 1, ct = wv & 1;
    const int arow = rt * 32 + (lane & 31), bcol = ct * 32 + (lane & 31), kh = (lane >> 5) * 8;
    const float bv = A.theta[OFF_B(LL) + bcol];
    const float *wo = A.theta + OFF_WOUT;
    double zr = 0.0, zq = 0.0, jr = 0.0, jq = 0.0;
    {
        const int tile0 = blockIdx.x * NNTILE;
        for (int qq = t; qq < NNTILE * 8; qq += NNT) {   // own features
            const int nd = qq >> 3, c8 = (qq & 7) * 8;
            const int i = tile0 + nd;
            u32x4 v = {0u, 0u, 0u, 0u};
            if (i < A.n) v = *(const u32x4 *)(Hin + (size_t)i * NNC + c8);
            *(u32x4 *)&sX[nd * LDX + c8] = v;
        }
        {   // aggregation: node t/4, channels (t%4)*16 .. +16, neighbours in CSR order
            const int nd = t >> 2, c0 = (t & 3) * 16;
            const int i = tile0 + nd;
            float acc[16];
            for (int c = 0; c < 16; ++c) acc[c] = 0.f;
            if (i < A.n) {
                const int p0 = A.rowptr[i] + 1, p1 = A.rowptr[i + 1];
                for (int pb = p0; pb < p1; pb += NGB) {
                    int jn[NGB];
                    float wn[NGB];
#pragma unroll
                    for (int u = 0; u < NGB; ++u) {
                        const int p = pb + u < p1 ? pb + u : p0;
                        jn[u] = A.colind[p];
                        wn[u] = A.w[p];
                    }
                    {
                        u32x4 va[NGB], vb[NGB];
#pragma unroll
                        for (int u = 0; u < NGB; ++u) {
                            va[u] = *(const u32x4 *)(Hin + (size_t)jn[u] * NNC + c0);
                            vb[u] = *(const u32x4 *)(Hin + (size_t)jn[u] * NNC + c0 + 8);
                        }
#pragma unroll
                        for (int u = 0; u < NGB; ++u) {
                            if (pb + u < p1) {
                                nn_acc8(acc, wn[u], va[u]);
                                nn_acc8(acc + 8, wn[u], vb[u]);
                            }
                        }
                    }
                }
            }
            unsigned int pk[8];
            for (int c = 0; c < 8; ++c)
                pk[c] = (unsigned int)f2bf(acc[2 * c]) | ((unsigned int)f2bf(acc[2 * c + 1]) << 16);
            const u32x4 v0 = {pk[0], pk[1], pk[2], pk[3]}, v1 = {pk[4], pk[5], pk[6], pk[7]};
            *(u32x4 *)&sX[nd * LDX + NNC + c0] = v0;
            *(u32x4 *)&sX[nd * LDX + NNC + c0 + 8] = v1;
            if (i < A.n && !A.st) {   // the aggregates are kept for the backward pass only
                *(u32x4 *)(Mout + (size_t)i * NNC + c0) = v0;
                *(u32x4 *)(Mout + (size_t)i * NNC + c0 + 8) = v1;
            }
        }
        __syncthreads();
        f32x16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < 8; ++ks) {
            const bf16x8 a = __builtin_bit_cast(bf16x8, *(const u32x4 *)&sX[arow * LDX + ks * 16 + kh]);
            const bf16x8 b = __builtin_bit_cast(bf16x8, *(const u32x4 *)&sW[bcol * LDX + ks * 16 + kh]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        __syncthreads();
        for (int v = 0; v < 16; ++v) {
            const int row = rt * 32 + (v & 3) + 4 * (lane >> 5) + 8 * (v >> 2);
            float h = acc[v] + bv;
            h = h > 0.f ? h : 0.f;
            sO[row * LDT + bcol] = f2bf(h);
        }
        __syncthreads();
        for (int qq = t; qq < NNTILE * 8; qq += NNT) {   // coalesced 16-byte stores of the tile
            const int nd = qq >> 3, c8 = (qq & 7) * 8;
            const int i = tile0 + nd;
            if (i < A.n) *(u32x4 *)(Hout + (size_t)i * NNC + c8) = *(const u32x4 *)&sO[nd * LDT + c8];
        }
        if (LL == 3) {   // y = h3 . w_out + b_out over channels in order
            const int nd = t >> 2, part = t & 3;
            float yv = 0.f;
            if (part < 2) {
                for (int c = part * 32; c < part * 32 + 32; ++c) yv = yv + bf2f(sO[nd * LDT + c]) * wo[c];
                sY[part][nd] = yv;
            }
            __syncthreads();
            if (t < NNTILE) {
                const int i = tile0 + t;
                if (i < A.n) {
                    const float ys = (sY[0][t] + sY[1][t]) + A.theta[OFF_BOUT];
                    A.y[i] = ys;
                    const double r_ = rin[i];
                    const double rt_ = A.isq[i] * r_;
                    const double z = A.isq[i] * (rt_ + s * (double)ys);
                    zout[i] = z;
                    zr += z * r_;
                    const double jz = A.isq[i] * A.isq[i] * r_;
                    jr += jz * r_;
                    if (q) {
                        zq += z * q[i];
                        jq += jz * q[i];
                    }
                }
            }
        }
        __syncthreads();
    }
    if (LL == 3 && A.st) {   // per-tile partials (ntiles of them) of z.r, z.q, r.D^-1 r, (D^-1 r).q
        nn_bsum3(zr, zq, jr);
        double d1 = 0.0, d2 = 0.0;
        nn_bsum3(jq, d1, d2);
        if (t == 0) {
            A.tpart[blockIdx.x] = zr;
            A.tpart[(size_t)A.ntiles + blockIdx.x] = zq;
            A.tpart[(size_t)2 * A.ntiles + blockIdx.x] = jr;
            A.tpart[(size_t)3 * A.ntiles + blockIdx.x] = jq;
        }
    }
}

// ------------------------------------------------------------------------------
// loss and backward
// ------------------------------------------------------------------------------
// A e with e = z - x, partials x.Ax and e.Ae
__global__ void __launch_bounds__(NNT) k_nn_loss(NNArgs A, const double *z, const double *x)
{
    double xax = 0.0, eae = 0.0, d2 = 0.0;
    for (int i = blockIdx.x * NNT + threadIdx.x; i < A.n; i += A.nbd * NNT) {
        double ax = 0.0, ae = 0.0;
        const int p0 = A.rowptr[i], p1 = A.rowptr[i + 1];
        for (int pb = p0; pb < p1; pb += NSG) {
            double xv[NSG], zv[NSG], vv[NSG];
#pragma unroll
            for (int u = 0; u < NSG; ++u) {
                const int p = pb + u < p1 ? pb + u : p0;
                const int j = A.colind[p];
                vv[u] = A.val[p];
                xv[u] = x[j];
                zv[u] = z[j];
            }
#pragma unroll
            for (int u = 0; u < NSG; ++u)
                if (pb + u < p1) {
                    ax = ax + vv[u] * xv[u];
                    ae = ae + vv[u] * (zv[u] - xv[u]);
                }
        }
        A.ae[i] = ae;
        xax += x[i] * ax;
        eae += (z[i] - x[i]) * ae;
    }
    nn_put3(A, SL_XAX, xax, SL_EAE, eae, -1, d2);
}

// dL/dy from dL/dz = 2 A e / (x.Ax) / B; dH3 = dy w_out; partial w_out, b_out
__global__ void __launch_bounds__(NNT) k_nn_out_bwd(NNArgs A, int batch)
{
    __shared__ float sg[NNT / 8][NNC + 1];
    double xax, eae, rtrt;
    nn_slots3(A.dpart, A.nbd, SL_XAX, SL_EAE, SL_RTRT, xax, eae, rtrt);
    if (blockIdx.x == 0 && threadIdx.x == 0) *A.loss += eae / xax / (double)batch;
    double s = sqrt(rtrt / (double)A.n);
    if (!(s > 0.0)) s = 1.0;
    const bf16_t *H3 = A.H + (size_t)3 * A.n * NNC;
    float *dH = A.dH;
    const float *wo = A.theta + OFF_WOUT;
    const int t = threadIdx.x, slot = t >> 3, c0 = (t & 7) * 8;
    float gw[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, gb = 0.f;
    for (int i = blockIdx.x * (NNT / 8) + slot; i < A.n; i += NNB * (NNT / 8)) {
        const double dz = 2.0 * A.ae[i] / xax / (double)batch;
        const float dy = (float)(dz * A.isq[i] * s);
        const u32x4 a = *(const u32x4 *)(H3 + (size_t)i * NNC + c0);
        const unsigned int wd[4] = {a.x, a.y, a.z, a.w};
        for (int c = 0; c < 4; ++c) {
            gw[2 * c] += bf2f((bf16_t)(wd[c] & 0xffffu)) * dy;
            gw[2 * c + 1] += bf2f((bf16_t)(wd[c] >> 16)) * dy;
        }
        for (int c = 0; c < 8; ++c) dH[(size_t)i * NNC + c0 + c] = dy * wo[c0 + c];
        if (c0 == 0) gb += dy;
    }
    for (int c = 0; c < 8; ++c) sg[slot][c0 + c] = gw[c];
    if (c0 == 0) sg[slot][NNC] = gb;
    __syncthreads();
    if (t <= NNC) {
        float a = 0.f;
        for (int k = 0; k < NNT / 8; ++k) a += sg[k][t];
        A.part[(size_t)blockIdx.x * NNP + (t < NNC ? OFF_WOUT + t : OFF_BOUT)] = a;
    }
}

// layer l backward, per tile: dpre = dH * (H_l > 0); dX = dpre W_l^T -> dXs, dXm
__global__ void __launch_bounds__(NNT) k_nn_layer_bwd_dx(NNArgs A, int l)
{
    __shared__ __attribute__((aligned(16))) bf16_t sD[NNTILE * LDT];
    __shared__ __attribute__((aligned(16))) bf16_t sW[128 * LDT];
    const int t = threadIdx.x;
    const int tile0 = blockIdx.x * NNTILE;
    const bf16_t *Hl = A.H + (size_t)l * A.n * NNC;
    const bf16_t *Wb = A.Wb + (size_t)(l - 1) * 128 * NNC;
    for (int q = t; q < 128 * 8; q += NNT) {   // W_l: 128 rows of 64
        const int row = q >> 3, c8 = (q & 7) * 8;
        *(u32x4 *)&sW[row * LDT + c8] = *(const u32x4 *)(Wb + row * NNC + c8);
    }
    for (int q = t; q < NNTILE * 8; q += NNT) {
        const int nd = q >> 3, c8 = (q & 7) * 8;
        const int i = tile0 + nd;
        unsigned int pk[4] = {0u, 0u, 0u, 0u};
        if (i < A.n) {
            const u32x4 hm = *(const u32x4 *)(Hl + (size_t)i * NNC + c8);
            const unsigned int hw[4] = {hm.x, hm.y, hm.z, hm.w};
            const float *g = A.dH + (size_t)i * NNC + c8;
            for (int c = 0; c < 4; ++c) {
                const float h0 = bf2f((bf16_t)(hw[c] & 0xffffu)), h1 = bf2f((bf16_t)(hw[c] >> 16));
                const float d0 = g[2 * c] * (h0 > 0.f ? 1.f : 0.f);
                const float d1 = g[2 * c + 1] * (h1 > 0.f ? 1.f : 0.f);
                pk[c] = (unsigned int)f2bf(d0) | ((unsigned int)f2bf(d1) << 16);
            }
        }
        const u32x4 v = {pk[0], pk[1], pk[2], pk[3]};
        *(u32x4 *)&sD[nd * LDT + c8] = v;
    }
    __syncthreads();
    const int wv = t >> 6, lane = t & 63;
    const int rt = wv >> 1;
    const int arow = rt * 32 + (lane & 31), kh = (lane >> 5) * 8;
    for (int h = 0; h < 2; ++h) {
        const int ft = (wv & 1) * 2 + h;   // feature block of 32
        const int bcol = ft * 32 + (lane & 31);
        f32x16 acc = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
        for (int ks = 0; ks < 4; ++ks) {
            const bf16x8 a = __builtin_bit_cast(bf16x8, *(const u32x4 *)&sD[arow * LDT + ks * 16 + kh]);
            const bf16x8 b = __builtin_bit_cast(bf16x8, *(const u32x4 *)&sW[bcol * LDT + ks * 16 + kh]);
            acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc, 0, 0, 0);
        }
        float *dst = bcol < NNC ? A.dXs : A.dXm;
        const int cc = bcol & (NNC - 1);
        for (int v = 0; v < 16; ++v) {
            const int row = rt * 32 + (v & 3) + 4 * (lane >> 5) + 8 * (v >> 2);
            const int i = tile0 + row;
            if (i < A.n) dst[(size_t)i * NNC + cc] = acc[v];
        }
    }
}

// layer l backward, grid-stride over tiles: partial dW_l = X^T dpre, db_l = sum dpre
__global__ void __launch_bounds__(NNT) k_nn_layer_bwd_dw(NNArgs A, int l)
{
    __shared__ __attribute__((aligned(16))) bf16_t sXt[128 * LDT];
    __shared__ __attribute__((aligned(16))) bf16_t sDt[NNC * LDT];
    __shared__ float sdb[4][NNC];
    const int t = threadIdx.x;
    const bf16_t *Hp = A.H + (size_t)(l - 1) * A.n * NNC;
    const bf16_t *Hl = A.H + (size_t)l * A.n * NNC;
    const bf16_t *Ml = A.M + (size_t)(l - 1) * A.n * NNC;
    const int wv = t >> 6, lane = t & 63, kh = (lane >> 5) * 8;
    f32x16 acc[2];
    for (int h = 0; h < 2; ++h)
        for (int v = 0; v < 16; ++v) acc[h][v] = 0.f;
    float db = 0.f;
    const int dc = t & 63, dg = t >> 6;   // bias: channel, 16-node group
    for (int tile = blockIdx.x; tile < A.ntiles; tile += NNB) {
        const int tile0 = tile * NNTILE;
        __syncthreads();
        for (int q = t; q < NNTILE * 16; q += NNT) {   // X^T: 128 features x 64 nodes
            const int nd = q & 63, f8 = (q >> 6) * 8;
            const int i = tile0 + nd;
            u32x4 v = {0u, 0u, 0u, 0u};
            if (i < A.n) v = f8 < NNC ? *(const u32x4 *)(Hp + (size_t)i * NNC + f8)
                                      : *(const u32x4 *)(Ml + (size_t)i * NNC + f8 - NNC);
            const unsigned int wd[4] = {v.x, v.y, v.z, v.w};
            for (int c = 0; c < 4; ++c) {
                sXt[(f8 + 2 * c) * LDT + nd] = (bf16_t)(wd[c] & 0xffffu);
                sXt[(f8 + 2 * c + 1) * LDT + nd] = (bf16_t)(wd[c] >> 16);
            }
        }
        for (int q = t; q < NNTILE * NNC; q += NNT) {   // dpre^T: 64 channels x 64 nodes
            const int c = q & 63, nd = q >> 6;
            const int i = tile0 + nd;
            float d = 0.f;
            if (i < A.n) {
                const float h = bf2f(Hl[(size_t)i * NNC + c]);
                d = A.dH[(size_t)i * NNC + c] * (h > 0.f ? 1.f : 0.f);
            }
            sDt[c * LDT + nd] = f2bf(d);
            if (c == dc) db += d;   // thread t holds channel t%64 for nodes t/64, +4, ...
        }
        __syncthreads();
        for (int h = 0; h < 2; ++h) {
            const int arow = wv * 32 + (lane & 31), bcol = h * 32 + (lane & 31);
            for (int ks = 0; ks < 4; ++ks) {
                const bf16x8 a = __builtin_bit_cast(bf16x8, *(const u32x4 *)&sXt[arow * LDT + ks * 16 + kh]);
                const bf16x8 b = __builtin_bit_cast(bf16x8, *(const u32x4 *)&sDt[bcol * LDT + ks * 16 + kh]);
                acc[h] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[h], 0, 0, 0);
            }
        }
    }
    float *P = A.part + (size_t)blockIdx.x * NNP;
    for (int h = 0; h < 2; ++h)
        for (int v = 0; v < 16; ++v) {
            const int f = wv * 32 + (v & 3) + 4 * (lane >> 5) + 8 * (v >> 2);
            const int c = h * 32 + (lane & 31);
            P[OFF_W(l) + f * NNC + c] = acc[h][v];
        }
    sdb[dg][dc] = db;
    __syncthreads();
    if (t < NNC) P[OFF_B(l) + t] = sdb[0][t] + sdb[1][t] + sdb[2][t] + sdb[3][t];
}

// dH_{l-1} = dXs + Ahat^T dXm
__global__ void __launch_bounds__(NNT) k_nn_agg_bwd(NNArgs A)
{
    const int q = blockIdx.x * NNT + threadIdx.x;
    if (q >= A.n * 8) return;
    const int i = q >> 3, c0 = (q & 7) * 8;
    float acc[8];
    for (int c = 0; c < 8; ++c) acc[c] = A.dXs[(size_t)i * NNC + c0 + c];
    float m[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int p = A.rowptr[i] + 1; p < A.rowptr[i + 1]; ++p) {
        const int j = A.colind[p];
        const float wv = A.wt[p];
        for (int c = 0; c < 8; ++c) m[c] = m[c] + wv * A.dXm[(size_t)j * NNC + c0 + c];
    }
    for (int c = 0; c < 8; ++c) A.dH[(size_t)i * NNC + c0 + c] = acc[c] + m[c];
}

// input layer backward: partial W0 = [sum u dpre0, sum dpre0]
__global__ void __launch_bounds__(NNT) k_nn_in_bwd(NNArgs A)
{
    __shared__ float sg[NNT / 8][2 * NNC];
    const int t = threadIdx.x, slot = t >> 3, c0 = (t & 7) * 8;
    float gu[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f}, g1[8] = {0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
    for (int i = blockIdx.x * (NNT / 8) + slot; i < A.n; i += NNB * (NNT / 8)) {
        const float u = A.u[i];
        for (int c = 0; c < 8; ++c) {
            const float h = bf2f(A.H[(size_t)i * NNC + c0 + c]);
            const float d = A.dH[(size_t)i * NNC + c0 + c] * (h > 0.f ? 1.f : 0.f);
            gu[c] += u * d;
            g1[c] += d;
        }
    }
    for (int c = 0; c < 8; ++c) { sg[slot][c0 + c] = gu[c]; sg[slot][NNC + c0 + c] = g1[c]; }
    __syncthreads();
    if (t < 2 * NNC) {
        float a = 0.f;
        for (int k = 0; k < NNT / 8; ++k) a += sg[k][t];
        A.part[(size_t)blockIdx.x * NNP + OFF_W0 + t] = a;
    }
}

// gradient = sum of the NNB partial rows (in block order); added to gacc unless first
__global__ void __launch_bounds__(NNT) k_nn_reduce(NNArgs A, float *gacc, int first)
{
    const int j = blockIdx.x * NNT + threadIdx.x;
    if (j >= NNP) return;
    float s = 0.f;
    for (int b = 0; b < NNB; ++b) s += A.part[(size_t)b * NNP + j];
    gacc[j] = first ? s : gacc[j] + s;
}

// bf16 copies of the layer weights in both layouts
__global__ void __launch_bounds__(NNT) k_nn_pack(const float *theta, bf16_t *Wf, bf16_t *Wb)
{
    const int q = blockIdx.x * NNT + threadIdx.x;
    if (q >= 3 * 128 * NNC) return;
    const int l = q / (128 * NNC), r = q % (128 * NNC);
    const int in = r / NNC, out = r % NNC;
    const bf16_t b = f2bf(theta[OFF_W(l + 1) + r]);
    Wb[(size_t)l * 128 * NNC + in * NNC + out] = b;
    Wf[(size_t)l * 128 * NNC + out * 128 + in] = b;
}

__global__ void __launch_bounds__(NNT) k_nn_adam(float *theta, const float *g, float *m, float *v, float lr, float b1,
                                                 float b2, float eps, float c1, float c2)
{
    const int j = blockIdx.x * NNT + threadIdx.x;
    if (j >= NNP) return;
    const float gj = g[j];
    const float mj = b1 * m[j] + (1.f - b1) * gj;
    const float vj = b2 * v[j] + (1.f - b2) * gj * gj;
    m[j] = mj;
    v[j] = vj;
    const float mh = mj / c1, vh = vj / c2;
    theta[j] = theta[j] - lr * mh / (sqrtf(vh) + eps);
}

// ------------------------------------------------------------------------------
// flexible CG around the network
// ------------------------------------------------------------------------------
struct FCGArgs {
    int n, nbd, maxiter, ntiles;
    const double *tpart;
    const int *rowptr, *colind;
    const double *val, *isq, *rhs;
    double *x, *r, *z, *p, *p1, *q, *dpart;
    FCGState *st;
};

__device__ double fcg_slot(const FCGArgs &F, int slot)
{
    double x = nn_strided_sum(F.dpart + (size_t)slot * F.nbd, F.nbd), d1 = 0.0, d2 = 0.0;
    nn_bsum3(x, d1, d2);
    return x;
}

__device__ __forceinline__ void fcg_put(const FCGArgs &F, int s0, double a, int s1, double b, int s2, double c)
{
    nn_bsum3(a, b, c);
    if (threadIdx.x == 0) {
        if (s0 >= 0) F.dpart[(size_t)s0 * F.nbd + blockIdx.x] = a;
        if (s1 >= 0) F.dpart[(size_t)s1 * F.nbd + blockIdx.x] = b;
        if (s2 >= 0) F.dpart[(size_t)s2 * F.nbd + blockIdx.x] = c;
    }
}

// r = b - A x; partials b.b, r.r, (D^-1/2 r)^2; state reset
__global__ void __launch_bounds__(NNT) k_fcg_init(FCGArgs F, double soltol, int maxiter)
{
    double bb = 0.0, rr = 0.0, tt = 0.0;
    for (int i = blockIdx.x * NNT + threadIdx.x; i < F.n; i += F.nbd * NNT) {
        double ax = 0.0;
        const int p0 = F.rowptr[i], p1 = F.rowptr[i + 1];
        for (int pb = p0; pb < p1; pb += NSG) {
            double xv[NSG], vv[NSG];
#pragma unroll
            for (int u = 0; u < NSG; ++u) {
                const int p = pb + u < p1 ? pb + u : p0;
                vv[u] = F.val[p];
                xv[u] = F.x[F.colind[p]];
            }
#pragma unroll
            for (int u = 0; u < NSG; ++u)
                if (pb + u < p1) ax = ax + vv[u] * xv[u];
        }
        const double r = F.rhs[i] - ax;
        F.r[i] = r;
        F.p1[i] = 0.0;
        bb += F.rhs[i] * F.rhs[i];
        rr += r * r;
        const double rt = F.isq[i] * r;
        tt += rt * rt;
    }
    fcg_put(F, SL_BB, bb, SL_RR, rr, SL_RTRT, tt);
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        FCGState s{};
        s.soltol = soltol;
        s.maxiter = maxiter;
        *F.st = s;
    }
}

// Direction and operator product of iteration k = st->iter, one kernel:
//   p_k = z_k + beta p_{k-1}, beta = -(z_k.q_{k-1}) / (p_{k-1}.q_{k-1}) (p_0 = z_0),
//   q_k = A p_k, partial p_k.q_k.
// If the network's z is not a descent direction (z.r <= 0), this iteration
// takes the diagonal preconditioner's z = D^-1 r instead; flexible CG admits a
// preconditioner that changes from one iteration to the next. p is double-
// buffered by parity (p_k in P[k&1]) so that neighbours' p_k can be formed from
// z and p_{k-1} while other blocks write p_k; p.q partials likewise. The
// parity k & 1 comes with the launch (every graph replay runs an even number
// of iterations from an even one), not from the state the previous kernel's
// lead block advanced.
__global__ void __launch_bounds__(NNT) k_fcg_pa(FCGArgs F, int par)
{
    FCGState *st = F.st;
    if (st->done) return;
    const int it = st->iter;
    const bool first = it == 0;
    double zr, zq, jr, jq, pq, u1, u2;
    nn_slots3(F.tpart, F.ntiles, 0, 1, 2, zr, zq, jr);
    nn_slots3(F.tpart, F.ntiles, 3, -1, -1, jq, u1, u2);
    nn_slots3(F.dpart, F.nbd, first ? SL_BB : (par ? SL_PQ : SL_PQ1), -1, -1, pq, u1, u2);
    const bool jac = !(zr > 0.0);
    if (jac) { zr = jr; zq = jq; }
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        st->zr = zr;
        if (jac) st->nneg += 1;
    }
    const double be = first ? 0.0 : -zq / pq;
    const double *pold = par ? F.p : F.p1;
    double *pnew = par ? F.p1 : F.p;
    double pqn = 0.0;
    for (int i = blockIdx.x * NNT + threadIdx.x; i < F.n; i += F.nbd * NNT) {
        double sq = 0.0;
        const int p0 = F.rowptr[i], p1 = F.rowptr[i + 1];
        for (int pb = p0; pb < p1; pb += NSG) {
            // column indices first, then each gather in a straight-line batch
            // (jac and first are uniform; branching per entry would serialise)
            int jv[NSG];
            double vv[NSG], zv[NSG], ov[NSG];
#pragma unroll
            for (int u = 0; u < NSG; ++u) {
                const int p = pb + u < p1 ? pb + u : p0;
                jv[u] = F.colind[p];
                vv[u] = F.val[p];
            }
            if (jac) {
#pragma unroll
                for (int u = 0; u < NSG; ++u) zv[u] = F.isq[jv[u]] * F.isq[jv[u]] * F.r[jv[u]];
            } else {
#pragma unroll
                for (int u = 0; u < NSG; ++u) zv[u] = F.z[jv[u]];
            }
            if (!first) {
#pragma unroll
                for (int u = 0; u < NSG; ++u) ov[u] = pold[jv[u]];
            }
#pragma unroll
            for (int u = 0; u < NSG; ++u)
                if (pb + u < p1) sq = sq + vv[u] * (first ? zv[u] : zv[u] + be * ov[u]);
        }
        const double zi = jac ? F.isq[i] * F.isq[i] * F.r[i] : F.z[i];
        const double pi = first ? zi : zi + be * pold[i];
        pnew[i] = pi;
        F.q[i] = sq;
        pqn += pi * sq;
    }
    fcg_put(F, par ? SL_PQ1 : SL_PQ, pqn, -1, 0.0, -1, 0.0);
}

// alpha = (z.r)/(p.q); x += alpha p; r -= alpha q; partials r.r, (D^-1/2 r)^2.
// par = k & 1 comes with the launch: the lead block advances st->iter at the
// end, which a block dispatched after it must not read as this iteration's.
__global__ void __launch_bounds__(NNT) k_fcg_b(FCGArgs F, int par)
{
    FCGState *st = F.st;
    if (st->done) return;
    const double zr = st->zr, pq = fcg_slot(F, par ? SL_PQ1 : SL_PQ);
    const double *pk = par ? F.p1 : F.p;
    const bool lead = blockIdx.x == 0 && threadIdx.x == 0;
    if (!(pq > 0.0) || !(zr > 0.0)) {   // breakdown: continue with plain CG
        if (lead) {
            st->nbreak += 1;
            st->done = 2;
        }
        return;
    }
    const double al = zr / pq;
    double rr = 0.0, tt = 0.0, d2 = 0.0;
    for (int i = blockIdx.x * NNT + threadIdx.x; i < F.n; i += F.nbd * NNT) {
        F.x[i] = F.x[i] + al * pk[i];
        const double r = F.r[i] - al * F.q[i];
        F.r[i] = r;
        rr += r * r;
        const double rt = F.isq[i] * r;
        tt += rt * rt;
    }
    fcg_put(F, SL_RR, rr, SL_RTRT, tt, -1, d2);
    if (lead) {
        st->iter += 1;
        if (st->iter >= st->maxiter) st->done = 3;
    }
}

// ==============================================================================
// host side
// ==============================================================================
static NNArgs nn_args(fgpu_context *c, FCGState *st)
{
    NNDev &d = *c->nn;
    NNArgs A;
    A.n = d.n; A.ntiles = d.ntiles; A.nbd = d.nbd;
    A.rowptr = c->ssh_rowptr; A.colind = c->ssh_colind; A.val = c->ssh_val;
    A.w = d.w; A.wt = d.wt; A.isq = d.isq; A.theta = d.theta;
    A.Wf = d.Wf; A.Wb = d.Wb; A.H = d.H; A.M = d.M;
    A.u = d.u; A.y = d.y; A.dH = d.dH; A.dXs = d.dXs; A.dXm = d.dXm; A.part = d.part;
    A.dpart = d.dpart; A.ae = d.ae; A.loss = d.loss; A.tpart = d.tpart;
    A.st = st;
    return A;
}

void fgpu_nn_release(fgpu_context *c)
{
    if (!c->nn) return;
    if (c->nn->fcg_graph) (void)hipGraphExecDestroy(c->nn->fcg_graph);
    delete c->nn;   // device buffers belong to the context's allocation list
    c->nn = nullptr;
}

// graph-dependent buffers, built from the assembled SSH matrix
static int nn_setup(fgpu_context *c)
{
    if (!c->ssh_ready) return fgpu_fail(FGPU_ERR_STATE, "SSH matrix not set");
    if (c->nn && c->nn->ready && c->nn->n == c->ssh_n) return FGPU_OK;
    if (c->nn && c->nn->ready) return fgpu_fail(FGPU_ERR_STATE, "SSH matrix shape changed");
    if (!c->nn) c->nn = new NNDev();
    NNDev &d = *c->nn;
    const int n = c->ssh_n, nnz = c->ssh_nnz;
    d.n = n;
    d.nnz = nnz;
    d.ntiles = (n + NNTILE - 1) / NNTILE;
    d.nbd = c->ssh_blocks;
    int rc;
    struct B { void **p; size_t bytes; } bufs[] = {
        {(void **)&d.w, sizeof(float) * nnz}, {(void **)&d.wt, sizeof(float) * nnz},
        {(void **)&d.isq, sizeof(double) * n}, {(void **)&d.theta, sizeof(float) * NNP},
        {(void **)&d.gacc, sizeof(float) * NNP}, {(void **)&d.am, sizeof(float) * NNP},
        {(void **)&d.av, sizeof(float) * NNP}, {(void **)&d.Wf, sizeof(bf16_t) * 3 * 128 * NNC},
        {(void **)&d.Wb, sizeof(bf16_t) * 3 * 128 * NNC}, {(void **)&d.H, sizeof(bf16_t) * 4 * (size_t)n * NNC},
        {(void **)&d.M, sizeof(bf16_t) * 3 * (size_t)n * NNC}, {(void **)&d.u, sizeof(float) * n},
        {(void **)&d.y, sizeof(float) * n}, {(void **)&d.dH, sizeof(float) * (size_t)n * NNC},
        {(void **)&d.dXs, sizeof(float) * (size_t)n * NNC}, {(void **)&d.dXm, sizeof(float) * (size_t)n * NNC},
        {(void **)&d.part, sizeof(float) * (size_t)NNB * NNP}, {(void **)&d.dpart, sizeof(double) * NSLOT * d.nbd},
        {(void **)&d.tpart, sizeof(double) * 4 * (size_t)d.ntiles},
        {(void **)&d.ae, sizeof(double) * n}, {(void **)&d.zs, sizeof(double) * n},
        {(void **)&d.loss, sizeof(double)}, {(void **)&d.st, sizeof(FCGState)}};
    for (auto &b : bufs)
        if ((rc = fgpu_alloc(c, b.p, b.bytes))) return rc;
    FGPU_HIP(hipMemset(d.dpart, 0, sizeof(double) * NSLOT * d.nbd));
    FGPU_HIP(hipMemset(d.st, 0, sizeof(FCGState)));
    // Ahat and its transpose per CSR entry
    std::vector<double> val(nnz);
    FGPU_HIP(hipStreamSynchronize(c->stream));
    FGPU_HIP(hipMemcpy(val.data(), c->ssh_val, sizeof(double) * nnz, hipMemcpyDeviceToHost));
    const std::vector<int> &rp = c->h_ssh_rowptr, &ci = c->h_ssh_colind;
    std::vector<double> isq(n);
    for (int i = 0; i < n; ++i) {
        const double di = val[rp[i]];
        if (!(di > 0.0)) return fgpu_fail(FGPU_ERR_STATE, "SSH matrix has a non-positive diagonal");
        isq[i] = 1.0 / std::sqrt(di);
    }
    std::vector<float> w(nnz, 0.f), wt(nnz, 0.f);
    for (int i = 0; i < n; ++i)
        for (int p = rp[i] + 1; p < rp[i + 1]; ++p) w[p] = (float)(val[p] * isq[i] * isq[ci[p]]);
    for (int i = 0; i < n; ++i)
        for (int p = rp[i] + 1; p < rp[i + 1]; ++p) {
            const int j = ci[p];
            int found = -1;
            for (int q = rp[j] + 1; q < rp[j + 1]; ++q)
                if (ci[q] == i) { found = q; break; }
            if (found < 0) return fgpu_fail(FGPU_ERR_STATE, "SSH matrix pattern is not symmetric");
            wt[p] = w[found];
        }
    FGPU_HIP(hipMemcpy(d.w, w.data(), sizeof(float) * nnz, hipMemcpyHostToDevice));
    FGPU_HIP(hipMemcpy(d.wt, wt.data(), sizeof(float) * nnz, hipMemcpyHostToDevice));
    FGPU_HIP(hipMemcpy(d.isq, isq.data(), sizeof(double) * n, hipMemcpyHostToDevice));
    d.ready = true;
    return FGPU_OK;
}

static int nn_pack(fgpu_context *c)
{
    NNDev &d = *c->nn;
    hipLaunchKernelGGL(k_nn_pack, dim3((3 * 128 * NNC + NNT - 1) / NNT), dim3(NNT), 0, c->stream, d.theta, d.Wf, d.Wb);
    FGPU_HIP(hipGetLastError());
    return FGPU_OK;
}

static int nn_ready(fgpu_context *c)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    FGPU_HIP(hipSetDevice(c->device));
    int rc = nn_setup(c);
    if (rc) return rc;
    if (!c->nn->weights) return fgpu_fail(FGPU_ERR_STATE, "network weights not set (fgpu_ssh_nn_init)");
    return FGPU_OK;
}

// forward of the network on rin -> zout (training mode: scale from k_nn_rms)
static int nn_forward(fgpu_context *c, const NNArgs &A, const double *rin, double *zout, const double *q, double soltol)
{
    const dim3 gt(A.ntiles), bb(NNT);
    {
        ProfScope ps(c, "k_nn_in");
        hipLaunchKernelGGL(k_nn_in, dim3(A.nbd), bb, 0, c->stream, A, rin, soltol);
    }
    {
        ProfScope ps(c, "k_nn_layer1");
        hipLaunchKernelGGL(k_nn_layer<1>, gt, bb, 0, c->stream, A, rin, zout, q, soltol);
    }
    {
        ProfScope ps(c, "k_nn_layer2");
        hipLaunchKernelGGL(k_nn_layer<2>, gt, bb, 0, c->stream, A, rin, zout, q, soltol);
    }
    {
        ProfScope ps(c, "k_nn_layer3");
        hipLaunchKernelGGL(k_nn_layer<3>, gt, bb, 0, c->stream, A, rin, zout, q, soltol);
    }
    FGPU_HIP(hipGetLastError());
    return FGPU_OK;
}

// forward + backward of one training sample; gradient into gacc (first: overwrite)
static int nn_sample(fgpu_context *c, const NNArgs &A, const double *b, const double *x, int batch, bool first)
{
    NNDev &d = *c->nn;
    const dim3 gb(A.nbd), bb(NNT), gt(A.ntiles), gp(NNB);
    hipLaunchKernelGGL(k_nn_rms, gb, bb, 0, c->stream, A, b);
    int rc = nn_forward(c, A, b, d.zs, nullptr, 0.0);
    if (rc) return rc;
    hipLaunchKernelGGL(k_nn_loss, gb, bb, 0, c->stream, A, d.zs, x);
    hipLaunchKernelGGL(k_nn_out_bwd, gp, bb, 0, c->stream, A, batch);
    for (int l = 3; l >= 1; --l) {
        hipLaunchKernelGGL(k_nn_layer_bwd_dx, gt, bb, 0, c->stream, A, l);
        hipLaunchKernelGGL(k_nn_layer_bwd_dw, gp, bb, 0, c->stream, A, l);
        hipLaunchKernelGGL(k_nn_agg_bwd, dim3((A.n * 8 + NNT - 1) / NNT), bb, 0, c->stream, A);
    }
    hipLaunchKernelGGL(k_nn_in_bwd, gp, bb, 0, c->stream, A);
    hipLaunchKernelGGL(k_nn_reduce, dim3((NNP + NNT - 1) / NNT), bb, 0, c->stream, A, d.gacc, first ? 1 : 0);
    FGPU_HIP(hipGetLastError());
    return FGPU_OK;
}

extern "C" int fgpu_ssh_nn_param_count(void) { return NNP; }

extern "C" int fgpu_ssh_nn_init(fgpu_context *c, unsigned long long seed)
{
    if (!c) return fgpu_fail(FGPU_ERR_ARG, "null context");
    FGPU_HIP(hipSetDevice(c->device));
    int rc = nn_setup(c);
    if (rc) return rc;
    // He-style initialisation of the hidden layers, output weights zero: the
    // untrained network is the diagonal preconditioner
    std::mt19937_64 gen(seed);
    std::normal_distribution<double> nd(0.0, 1.0);
    std::vector<float> th(NNP, 0.f);
    for (int k = 0; k < 2 * NNC; ++k) th[OFF_W0 + k] = (float)nd(gen);
    const double sd = std::sqrt(2.0 / 128.0);
    for (int l = 1; l <= 3; ++l)
        for (int k = 0; k < 128 * NNC; ++k) th[OFF_W(l) + k] = (float)(sd * nd(gen));
    NNDev &d = *c->nn;
    FGPU_HIP(hipMemcpy(d.theta, th.data(), sizeof(float) * NNP, hipMemcpyHostToDevice));
    FGPU_HIP(hipMemset(d.am, 0, sizeof(float) * NNP));
    FGPU_HIP(hipMemset(d.av, 0, sizeof(float) * NNP));
    d.adam_t = 0;
    d.weights = true;
    rc = nn_pack(c);
    if (rc) return rc;
    FGPU_HIP(hipStreamSynchronize(c->stream));
    return FGPU_OK;
}

extern "C" int fgpu_ssh_nn_set_weights(fgpu_context *c, const float *theta, int np)
{
    if (!c || !theta) return fgpu_fail(FGPU_ERR_ARG, "null argument");
    if (np != NNP) return fgpu_fail(FGPU_ERR_ARG, "expected " + std::to_string(NNP) + " parameters");
    FGPU_HIP(hipSetDevice(c->device));
    int rc = nn_setup(c);
    if (rc) return rc;
    NNDev &d = *c->nn;
    FGPU_HIP(hipMemcpy(d.theta, theta, sizeof(float) * NNP, hipMemcpyHostToDevice));
    FGPU_HIP(hipMemset(d.am, 0, sizeof(float) * NNP));
    FGPU_HIP(hipMemset(d.av, 0, sizeof(float) * NNP));
    d.adam_t = 0;
    d.weights = true;
    rc = nn_pack(c);
    if (rc) return rc;
    FGPU_HIP(hipStreamSynchronize(c->stream));
    return FGPU_OK;
}

extern "C" int fgpu_ssh_nn_get_weights(fgpu_context *c, float *theta, int np)
{
    int rc = nn_ready(c);
    if (rc) return rc;
    if (!theta || np != NNP) return fgpu_fail(FGPU_ERR_ARG, "bad parameter buffer");
    FGPU_HIP(hipStreamSynchronize(c->stream));
    FGPU_HIP(hipMemcpy(theta, c->nn->theta, sizeof(float) * NNP, hipMemcpyDeviceToHost));
    return FGPU_OK;
}

extern "C" int fgpu_ssh_nn_apply(fgpu_context *c, const double *r, double *z)
{
    int rc = nn_ready(c);
    if (rc) return rc;
    if (!r || !z) return fgpu_fail(FGPU_ERR_ARG, "null vector");
    NNDev &d = *c->nn;
    const size_t nb = sizeof(double) * d.n;
    FGPU_HIP(hipMemcpyAsync(d.ae, r, nb, hipMemcpyHostToDevice, c->stream));
    NNArgs A = nn_args(c, nullptr);
    hipLaunchKernelGGL(k_nn_rms, dim3(A.nbd), dim3(NNT), 0, c->stream, A, d.ae);
    if ((rc = nn_forward(c, A, d.ae, d.zs, nullptr, 0.0))) return rc;
    FGPU_HIP(hipMemcpyAsync(z, d.zs, nb, hipMemcpyDeviceToHost, c->stream));
    FGPU_HIP(hipStreamSynchronize(c->stream));
    return FGPU_OK;
}

static int nn_upload_samples(fgpu_context *c, const double *b, const double *x, int nsamp)
{
    NNDev &d = *c->nn;
    if (nsamp > d.nsamp_cap) {
        int rc;
        if ((rc = fgpu_alloc(c, (void **)&d.samp_b, sizeof(double) * (size_t)nsamp * d.n))) return rc;
        if ((rc = fgpu_alloc(c, (void **)&d.samp_x, sizeof(double) * (size_t)nsamp * d.n))) return rc;
        d.nsamp_cap = nsamp;
    }
    const size_t nb = sizeof(double) * (size_t)nsamp * d.n;
    FGPU_HIP(hipMemcpy(d.samp_b, b, nb, hipMemcpyHostToDevice));
    FGPU_HIP(hipMemcpy(d.samp_x, x, nb, hipMemcpyHostToDevice));
    return FGPU_OK;
}

extern "C" int fgpu_ssh_nn_loss_grad(fgpu_context *c, const double *b, const double *x, int nsamp, double *loss,
                                     float *grad)
{
    int rc = nn_ready(c);
    if (rc) return rc;
    if (!b || !x || nsamp < 1) return fgpu_fail(FGPU_ERR_ARG, "bad samples");
    if ((rc = nn_upload_samples(c, b, x, nsamp))) return rc;
    NNDev &d = *c->nn;
    NNArgs A = nn_args(c, nullptr);
    FGPU_HIP(hipMemsetAsync(d.loss, 0, sizeof(double), c->stream));
    for (int s = 0; s < nsamp; ++s)
        if ((rc = nn_sample(c, A, d.samp_b + (size_t)s * d.n, d.samp_x + (size_t)s * d.n, nsamp, s == 0))) return rc;
    FGPU_HIP(hipStreamSynchronize(c->stream));
    if (loss) FGPU_HIP(hipMemcpy(loss, d.loss, sizeof(double), hipMemcpyDeviceToHost));
    if (grad) FGPU_HIP(hipMemcpy(grad, d.gacc, sizeof(float) * NNP, hipMemcpyDeviceToHost));
    return FGPU_OK;
}

extern "C" int fgpu_ssh_nn_train(fgpu_context *c, const double *b, const double *x, int nsamp, int epochs, int batch,
                                 double lr, double *loss_per_epoch)
{
    int rc = nn_ready(c);
    if (rc) return rc;
    if (!b || !x || nsamp < 1 || epochs < 1 || batch < 1) return fgpu_fail(FGPU_ERR_ARG, "bad training arguments");
    if ((rc = nn_upload_samples(c, b, x, nsamp))) return rc;
    NNDev &d = *c->nn;
    NNArgs A = nn_args(c, nullptr);
    const float b1 = 0.9f, b2 = 0.999f, eps = 1e-8f;
    const int nbatch = (nsamp + batch - 1) / batch;
    for (int ep = 0; ep < epochs; ++ep) {
        FGPU_HIP(hipMemsetAsync(d.loss, 0, sizeof(double), c->stream));
        for (int k = 0; k < nbatch; ++k) {
            const int s0 = k * batch, s1 = std::min(nsamp, s0 + batch);
            for (int s = s0; s < s1; ++s)
                if ((rc = nn_sample(c, A, d.samp_b + (size_t)s * d.n, d.samp_x + (size_t)s * d.n, (s1 - s0) * nbatch,
                                    s == s0)))
                    return rc;
            d.adam_t += 1;
            const float c1 = 1.f - std::pow(b1, (float)d.adam_t), c2 = 1.f - std::pow(b2, (float)d.adam_t);
            // the batch gradient is a mean over the epoch's batches; rescale to the batch mean
            hipLaunchKernelGGL(k_nn_adam, dim3((NNP + NNT - 1) / NNT), dim3(NNT), 0, c->stream, d.theta, d.gacc, d.am,
                               d.av, (float)lr, b1, b2, eps, c1, c2);
            if ((rc = nn_pack(c))) return rc;
        }
        FGPU_HIP(hipStreamSynchronize(c->stream));
        if (loss_per_epoch) FGPU_HIP(hipMemcpy(&loss_per_epoch[ep], d.loss, sizeof(double), hipMemcpyDeviceToHost));
    }
    if (d.fcg_graph) { (void)hipGraphExecDestroy(d.fcg_graph); d.fcg_graph = nullptr; }
    return FGPU_OK;
}

// ---- flexible CG ---------------------------------------------------------------
#define FCG_GRAPH_ITERS 8

static FCGArgs fcg_args(fgpu_context *c, int maxiter)
{
    NNDev &d = *c->nn;
    FCGArgs F;
    F.n = d.n; F.nbd = d.nbd; F.maxiter = maxiter; F.ntiles = d.ntiles; F.tpart = d.tpart;
    F.rowptr = c->ssh_rowptr; F.colind = c->ssh_colind; F.val = c->ssh_val; F.isq = d.isq; F.rhs = c->ssh_rhs;
    F.x = c->ssh_x; F.r = c->ssh_r; F.z = c->ssh_z; F.p = c->ssh_p; F.p1 = c->ssh_p1; F.q = c->ssh_q;
    F.dpart = d.dpart;
    F.st = d.st;
    return F;
}

// iteration i of a group of FCG_GRAPH_ITERS (an even number), parity i & 1
static int fcg_enqueue_iter(fgpu_context *c, const FCGArgs &F, const NNArgs &A, double soltol, int i)
{
    const dim3 gb(F.nbd), bb(NNT);
    {
        ProfScope ps(c, "k_fcg_pa");
        hipLaunchKernelGGL(k_fcg_pa, gb, bb, 0, c->stream, F, i & 1);
    }
    {
        ProfScope ps(c, "k_fcg_b");
        hipLaunchKernelGGL(k_fcg_b, gb, bb, 0, c->stream, F, i & 1);
    }
    return nn_forward(c, A, F.r, F.z, F.q, soltol);
}

// the graph of FCG_GRAPH_ITERS iterations, captured once: the tolerance and
// the iteration cap are read from the device state k_fcg_init writes
static int fcg_graph(fgpu_context *c, const FCGArgs &F, const NNArgs &A, double soltol)
{
    NNDev &d = *c->nn;
    if (d.fcg_graph) return FGPU_OK;
    hipGraph_t g;
    FGPU_HIP(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
    int rc = FGPU_OK;
    for (int i = 0; i < FCG_GRAPH_ITERS && !rc; ++i) rc = fcg_enqueue_iter(c, F, A, soltol, i);
    hipError_t e = hipStreamEndCapture(c->stream, &g);
    if (rc) return rc;
    if (e != hipSuccess) return fgpu_fail(FGPU_ERR_HIP, std::string("hipStreamEndCapture: ") + hipGetErrorString(e));
    e = hipGraphInstantiate(&d.fcg_graph, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return fgpu_fail(FGPU_ERR_HIP, std::string("hipGraphInstantiate: ") + hipGetErrorString(e));
    return FGPU_OK;
}

struct FcgResult {
    int iters, done, nbreak, nneg, fb_iters, fb_done;
    double rr, rtol;
};

// learned FCG from ssh_x; on breakdown or at the cap, plain Jacobi CG from its x
static int fcg_run(fgpu_context *c, double soltol, int maxiter, int fb_maxiter, FcgResult *res)
{
    NNDev &d = *c->nn;
    FCGArgs F = fcg_args(c, maxiter);
    NNArgs A = nn_args(c, d.st);
    int rc = FGPU_OK;
    if (!c->prof && (rc = fcg_graph(c, F, A, soltol))) return rc;
    hipLaunchKernelGGL(k_fcg_init, dim3(F.nbd), dim3(NNT), 0, c->stream, F, soltol, maxiter);
    if ((rc = nn_forward(c, A, F.r, F.z, nullptr, soltol))) return rc;
    if (c->prof) {   // per-kernel timing: iterations launched one by one
        FCGState hs;
        for (int it = 0; it < maxiter + 1; it += FCG_GRAPH_ITERS) {
            for (int i = 0; i < FCG_GRAPH_ITERS && !rc; ++i) rc = fcg_enqueue_iter(c, F, A, soltol, i);
            if (rc) return rc;
            FGPU_HIP(hipMemcpyAsync(&hs, d.st, sizeof(FCGState), hipMemcpyDeviceToHost, c->stream));
            FGPU_HIP(hipStreamSynchronize(c->stream));
            if (hs.done) break;
        }
        if ((rc = fgpu_prof_flush(c))) return rc;
    }
    const int max_graphs = c->prof ? 0 : maxiter / FCG_GRAPH_ITERS + 2;
    int batch = d.fcg_last_iters > 0 ? (d.fcg_last_iters + FCG_GRAPH_ITERS - 1) / FCG_GRAPH_ITERS : 2;
    int launched = 0;
    FCGState hs;
    while (!c->prof) {
        for (int i = 0; i < batch && launched < max_graphs; ++i, ++launched)
            FGPU_HIP(hipGraphLaunch(d.fcg_graph, c->stream));
        FGPU_HIP(hipMemcpyAsync(&hs, d.st, sizeof(FCGState), hipMemcpyDeviceToHost, c->stream));
        FGPU_HIP(hipStreamSynchronize(c->stream));
        if (hs.done || launched >= max_graphs) break;
        batch = 1;
    }
    FGPU_HIP(hipMemcpy(&hs, d.st, sizeof(FCGState), hipMemcpyDeviceToHost));
    if (!hs.done) return fgpu_fail(FGPU_ERR_STATE, "FCG iteration graph did not terminate");
    d.fcg_last_iters = hs.iter;
    res->iters = hs.iter;
    res->done = hs.done;
    res->nbreak = hs.nbreak;
    res->nneg = hs.nneg;
    res->rr = hs.rr_last;
    res->rtol = hs.rtol;
    res->fb_iters = 0;
    res->fb_done = 0;
    if (hs.done != 1 && fb_maxiter > 0) {
        FgpuCgSummary s;
        if ((rc = fgpu_ssh_cg_from_x(c, soltol, fb_maxiter, 2, &s))) return rc;
        res->fb_iters = s.iters;
        res->fb_done = s.done;
        res->rr = s.rr_last;
    }
    return FGPU_OK;
}

static void fcg_stats(fgpu_context *c, const FcgResult &r, double ms, fgpu_cg_stats *st, int *fb_iters)
{
    if (st) {
        st->iters_last = r.iters + r.fb_iters;
        st->converged = r.done == 1 || r.fb_done == 1;
        st->nbreakdown = r.nbreak;
        st->nnegrz = r.nneg;
        st->resid_last = std::sqrt(r.rr / (double)c->ssh_n);
        st->rtol_last = r.rtol;
        st->ms = ms;
    }
    if (fb_iters) *fb_iters = r.fb_iters;
}

static int fcg_ready(fgpu_context *c, double soltol, int maxiter)
{
    int rc = nn_ready(c);
    if (rc) return rc;
    if (!c->ssh_pr_ready) return fgpu_fail(FGPU_ERR_STATE, "preconditioner not built (fgpu_ssh_solve_preconditioner)");
    if (maxiter < 1 || !(soltol > 0.0)) return fgpu_fail(FGPU_ERR_ARG, "soltol must be > 0 and maxiter >= 1");
    return FGPU_OK;
}

extern "C" int fgpu_ssh_solve_learned(fgpu_context *c, double *x, const double *rhs, double soltol, int maxiter,
                                      int fallback_maxiter, fgpu_cg_stats *stats, int *fallback_iters)
{
    int rc = fcg_ready(c, soltol, maxiter);
    if (rc) return rc;
    if (!x || !rhs) return fgpu_fail(FGPU_ERR_ARG, "null vector");
    const size_t nb = sizeof(double) * c->ssh_n;
    FGPU_HIP(hipMemcpyAsync(c->ssh_x, x, nb, hipMemcpyHostToDevice, c->stream));
    FGPU_HIP(hipMemcpyAsync(c->ssh_rhs, rhs, nb, hipMemcpyHostToDevice, c->stream));
    hipEvent_t e0, e1;
    FGPU_HIP(hipEventCreate(&e0));
    FGPU_HIP(hipEventCreate(&e1));
    FGPU_HIP(hipEventRecord(e0, c->stream));
    FcgResult r{};
    rc = fcg_run(c, soltol, maxiter, fallback_maxiter, &r);
    FGPU_HIP(hipEventRecord(e1, c->stream));
    FGPU_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    FGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    FGPU_HIP(hipMemcpy(x, c->ssh_x, nb, hipMemcpyDeviceToHost));
    fcg_stats(c, r, ms, stats, fallback_iters);
    return FGPU_OK;
}

__global__ void k_nn_copy(double *dst, const double *src, int n)
{
    const int i = blockIdx.x * NNT + threadIdx.x;
    if (i < n) dst[i] = src[i];
}

extern "C" int fgpu_ssh_solve_learned_resident(fgpu_context *c, double soltol, int maxiter, int fallback_maxiter,
                                               int nrep, fgpu_cg_stats *stats, int *fallback_iters)
{
    int rc = fcg_ready(c, soltol, maxiter);
    if (rc) return rc;
    if (nrep < 1) return fgpu_fail(FGPU_ERR_ARG, "nrep < 1");
    hipEvent_t e0, e1;
    FGPU_HIP(hipEventCreate(&e0));
    FGPU_HIP(hipEventCreate(&e1));
    float total = 0.f;
    FcgResult r{};
    for (int i = 0; i < nrep && !rc; ++i) {
        hipLaunchKernelGGL(k_nn_copy, dim3((c->ssh_n + NNT - 1) / NNT), dim3(NNT), 0, c->stream, c->ssh_x, c->ssh_x0,
                           c->ssh_n);
        FGPU_HIP(hipGetLastError());
        FGPU_HIP(hipEventRecord(e0, c->stream));
        rc = fcg_run(c, soltol, maxiter, fallback_maxiter, &r);
        FGPU_HIP(hipEventRecord(e1, c->stream));
        FGPU_HIP(hipEventSynchronize(e1));
        float ms = 0.f;
        FGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
        total += ms;
    }
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc) return rc;
    fcg_stats(c, r, total / nrep, stats, fallback_iters);
    return FGPU_OK;
}
