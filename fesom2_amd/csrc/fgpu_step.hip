// Fused, batched tracer step for gfx950: five launches per step for all tracers.
//
//   K_G  node pass   area-weighted mean element gradient of T^n at the levels
//                    where an edge needs it (fill_up_dn_grad's non-shared levels)
//   K_A  edge pass   LO upwind flux (T^n) and antidiffusive flux HO(T^AB) - LO
//                    (MFCT/UPW1); the up/downwind element gradients and T^AB are
//                    evaluated on the fly
//   K_B  node pass   LO solution and antidiffusive vertical flux (QR4C/UPW1)
//   K_C  node pass   FCT bounds over the 1-ring (edge gather), P+/P-, R+/R-
//   K_D  node pass   limited vertical and horizontal fluxes, flux2dtracer, T*,
//                    implicit vertical diffusion (Thomas), salinity clamp,
//                    valuesold rotation
//
// K_B..K_D are separated by the three global dependencies of FCT (LO of the
// neighbours, R+/R- of the neighbours, limited edge fluxes). One block owns one
// column (an edge or a node) and every tracer; lanes run over levels, so a
// column's topology and geometry are wave-uniform and its field accesses are
// contiguous column segments. The node kernels load the metadata of a node's
// edges (or elements) once, one edge per lane, and hand it to the gather loops
// with v_readlane, so the per-tracer loops carry no dependent scalar loads.
// Blocks are mapped to columns XCD by XCD (workgroups are dispatched to the
// XCDs round-robin), so neighbouring columns share an L2.
//
// Arithmetic is the same sequence of operations as the per-function kernels
// and the reference (no FMA contraction); edge->node sums run over the node's
// edges in ascending edge order. Results are bit-identical to a serial run of
// the reference's loops.
#include <algorithm>

#include "fgpu_internal.h"
#include "fgpu_vert.h"

#define NG 8    // edges per gather batch in K_B and K_C

__device__ __forceinline__ double dmx(double a, double b) { return a > b ? a : b; }
__device__ __forceinline__ double dmn(double a, double b) { return a < b ? a : b; }

// ordering of LDS accesses between the lanes of one wave: prior LDS writes
// complete before later accesses, and the compiler keeps the order
__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup");
}

__device__ __forceinline__ int rl(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }
__device__ __forceinline__ double rld(double v, int lane)
{
    const long long b = __double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(b & 0xffffffffLL), lane);
    const int hi = __builtin_amdgcn_readlane((int)(b >> 32), lane);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}

// position of this block's column in the processing order (the node or edge
// is node_ord / edge_ord of it), XCD-contiguous: blocks b, b+8, ... run on one
// XCD and take consecutive positions
__device__ __forceinline__ int xcd_column(int ncol)
{
    const int per = (int)(gridDim.x / NXCD);
    return (int)(blockIdx.x % NXCD) * per + (int)(blockIdx.x / NXCD);
}

__device__ __forceinline__ double upw1_h(double t1, double t2, double vflux, double fin)
{
    return -0.5 * (t1 * (vflux + fabs(vflux)) + t2 * (vflux - fabs(vflux))) - fin;
}

// T^AB from values and valuesold (oce_tracer_mod.F90:44-59)
__device__ __forceinline__ double t_ab(int abo, const double *T, const double *Told, size_t i)
{
    if (abo == 2) return -(0.5 + FGPU_EPS_AB) * Told[i] + (1.5 + FGPU_EPS_AB) * T[i];
    const double a = 5.0 * Told[2 * i + 1] - 16.0 * Told[2 * i] + 23.0 * T[i];
    return a / 12.0;
}

__device__ __forceinline__ double seg_vflux(int seg, double u1, double v1, double u2, double v2, double h1,
                                            double h2, double dX1, double dY1, double dX2, double dY2)
{
    if (seg == 0 || seg == 3) return (-v1 * dX1 + u1 * dY1) * h1;
    if (seg == 1 || seg == 4) return (v2 * dX2 - u2 * dY2) * h2;
    return (-v1 * dX1 + u1 * dY1) * h1 + (v2 * dX2 - u2 * dY2) * h2;
}

// limited antidiffusive edge flux as seen from a node (b3 horizontal,
// oce_adv_tra_fct.F90:483-497): R of this node (rpn, rmn) and of the edge's
// other node (rpo, rmo); q = 2e + side, side 1 when this node is the edge's
// second node
__device__ __forceinline__ double limit_edge_flux(double fl, int q, double rpn, double rmn, double rpo, double rmo)
{
    const double rp1 = (q & 1) ? rpo : rpn, rm1 = (q & 1) ? rmo : rmn;
    const double rp2 = (q & 1) ? rpn : rpo, rm2 = (q & 1) ? rmn : rmo;
    double ae = 1.0;
    if (fl >= 0.0) {
        ae = dmn(ae, rp1);
        ae = dmn(ae, rm2);
    } else {
        ae = dmn(ae, rm1);
        ae = dmn(ae, rp2);
    }
    return ae * fl;
}

// metadata of a node's edges, one edge per lane (lane < deg)
struct NodeEdges {
    int deg;
    int q;          // 2*e + side
    int other;      // the edge's other node
    int nu12, nl12; // edge scatter range (max/min rule)
    int nu1, nl1, nu2, nl2;
};

__device__ __forceinline__ NodeEdges load_node_edges(const DevMesh &m, int n)
{
    NodeEdges E;
    const int4 nr = *(const int4 *)(m.node_rec + 4 * (size_t)n);
    const int p0 = nr.x;
    E.deg = nr.y;
    const int j = threadIdx.x & 63;
    E.q = 0; E.other = 0; E.nu12 = 1; E.nl12 = 0; E.nu1 = 1; E.nl1 = 0; E.nu2 = 0; E.nl2 = 0;
    if (j < E.deg) {
        const int4 a = *(const int4 *)(m.ne_rec + 8 * (size_t)(p0 + j));
        const int4 b = *(const int4 *)(m.ne_rec + 8 * (size_t)(p0 + j) + 4);
        E.q = a.x;
        E.other = a.y;
        E.nu12 = a.z;
        E.nl12 = a.w;
        E.nl1 = b.x;
        E.nu1 = b.y;
        E.nl2 = b.z;
        E.nu2 = b.w;
    }
    return E;
}

// ------------------------------------------------------------------------------
// K_G: mean gradient around a node (oce_muscl_adv.F90:388-401, 443-520) for
// levels kneed(n)..nlevels_nod2D-1
// ------------------------------------------------------------------------------
template <int LB>
__global__ void __launch_bounds__(LB) kG_node(DevMesh m, DevTracers t, const int *kneed, int tpb)
{
    const int jn = xcd_column(m.N2);
    if (jn >= m.N2) return;
    const int n = m.node_ord[jn];
    const int k0 = kneed[n], nzmax = m.nlevn[n];
    if (k0 > nzmax - 1) return;
    const int L = m.L;
    const int k = threadIdx.x + 1;
    const int num = m.nie_num[n];
    const int j = threadIdx.x & 63;
    int lo = 1, hi = 0, a0 = 0, a1 = 0, a2 = 0;
    double ar = 0.0, g0 = 0.0, g1 = 0.0, g2 = 0.0, g3 = 0.0, g4 = 0.0, g5 = 0.0;
    if (j < num) {
        const int el = m.nie[(size_t)n * m.kmax + j];
        lo = m.ulev[el];
        hi = m.nlev[el] - 1;
        ar = m.elem_area[el];
        a0 = m.elem_nodes[3 * el]; a1 = m.elem_nodes[3 * el + 1]; a2 = m.elem_nodes[3 * el + 2];
        const double *gs = m.gsca + 6 * (size_t)el;
        g0 = gs[0]; g1 = gs[1]; g2 = gs[2]; g3 = gs[3]; g4 = gs[4]; g5 = gs[5];
    }
    const bool act = (k >= k0 && k <= nzmax - 1);
    const int trb = blockIdx.y * tpb, tre = min(t.ntr, trb + tpb);
    for (int tr = trb; tr < tre; ++tr) {
        if (t.hor[tr] == HOR_UPW1) continue;   // node means feed MFCT and MUSCL
        const double *T = t.T + t.sN * tr;
        double tvol = 0.0, tx = 0.0, ty = 0.0;
        for (int i = 0; i < num; ++i) {
            const int elo = rl(lo, i), ehi = rl(hi, i);
            const int b0 = rl(a0, i), b1 = rl(a1, i), b2 = rl(a2, i);
            const double a = rld(ar, i);
            const double s0 = rld(g0, i), s1 = rld(g1, i), s2 = rld(g2, i);
            const double s3 = rld(g3, i), s4 = rld(g4, i), s5 = rld(g5, i);
            if (act && k >= elo && k <= ehi) {
                const double t0 = T[(size_t)b0 * L + k - 1], t1 = T[(size_t)b1 * L + k - 1],
                             t2 = T[(size_t)b2 * L + k - 1];
                const double ex = s0 * t0 + s1 * t1 + s2 * t2;
                const double ey = s3 * t0 + s4 * t1 + s5 * t2;
                tvol = tvol + a;
                tx = tx + ex * a;
                ty = ty + ey * a;
            }
        }
        if (act) {
            double *o = t.gmean + t.sG * tr + ((size_t)n * L + k - 1) * 2;
            o[0] = tx / tvol;
            o[1] = ty / tvol;
        }
    }
}

// ------------------------------------------------------------------------------
// K_A: edge pass
// ------------------------------------------------------------------------------
template <int LB>
__global__ void __launch_bounds__(LB) __attribute__((amdgpu_waves_per_eu(6, 8))) kA_edge(DevMesh m, DevDyn d, DevTracers t, int tpb)
{
    const int je = xcd_column(m.E2);
    if (je >= m.E2) return;
    const int e = m.edge_ord[je];
    const int k = threadIdx.x + 1;
    const int L = m.L;
    const int *ri = m.erec_i + (size_t)e * ERI;
    const double *rd = m.erec_d + (size_t)e * ERD;
    const int nl1 = ri[ER_NL1], nu1 = ri[ER_NU1], nl2 = ri[ER_NL2], nu2 = ri[ER_NU2];
    const int nl12 = min(nl1, nl2), nu12 = max(nu1, nu2);
    bool in[5];
    in[0] = (k >= nu1 && k <= nu12 - 1);
    in[1] = (nu2 > 0 && k >= nu2 && k <= nu12 - 1);
    in[2] = (k >= nu12 && k <= nl12);
    in[3] = (k >= nl12 + 1 && k <= nl1);
    in[4] = (k >= nl12 + 1 && k <= nl2);
    if (!(in[0] || in[1] || in[2] || in[3] || in[4])) return;
    const int n1 = ri[ER_N1], n2 = ri[ER_N2], el1 = ri[ER_EL1], el2 = ri[ER_EL2];
    const double dX1 = rd[ER_DX1], dY1 = rd[ER_DY1], dX2 = rd[ER_DX2], dY2 = rd[ER_DY2];
    const double a = rd[ER_A], edx = rd[ER_EDX], edy = rd[ER_EDY];
    const size_t q1 = (size_t)el1 * L + k - 1;
    const double u1 = d.uv[2 * q1], v1 = d.uv[2 * q1 + 1], h1 = m.helem[q1];
    double u2 = 0.0, v2 = 0.0, h2 = 0.0;
    if (el2 >= 0 && (in[1] || in[2] || in[4])) {
        const size_t q2 = (size_t)el2 * L + k - 1;
        u2 = d.uv[2 * q2];
        v2 = d.uv[2 * q2 + 1];
        h2 = m.helem[q2];
    }
    // at most two segments meet at one level, in this order: A then B, or C, or
    // D then E (A..E as in oce_adv_tra_hor.F90:166-249)
    const int s1 = in[0] ? 0 : (in[2] ? 2 : (in[3] ? 3 : -1));
    const int s2 = in[1] ? 1 : (in[4] ? 4 : -1);
    const double vf1 = s1 >= 0 ? seg_vflux(s1, u1, v1, u2, v2, h1, h2, dX1, dY1, dX2, dY2) : 0.0;
    const double vf2 = s2 >= 0 ? seg_vflux(s2, u1, v1, u2, v2, h1, h2, dX1, dY1, dX2, dY2) : 0.0;
    // gradient rule of fill_up_dn_grad at this level (oce_muscl_adv.F90:375-520);
    // writes in the reference's order, a later write wins
    bool shared = false, mean1 = false, mean2 = false;
    const int u1l = ri[ER_U1], u2l = ri[ER_U2], nl1n = ri[ER_NL1N], nl2n = ri[ER_NL2N];
    if (ri[ER_RULE]) {
        const int nzmin = ri[ER_NZMIN], nzmax = ri[ER_NZMAX];
        mean1 = (k >= u1l && k <= nzmin - 1);
        mean2 = (k >= u2l && k <= nzmin - 1);
        if (k >= nzmin && k <= nzmax - 1) { shared = true; mean1 = false; mean2 = false; }
        if (k >= nzmax && k <= nl1n - 1) { mean1 = true; shared = false; }
        if (k >= nzmax && k <= nl2n - 1) { mean2 = true; shared = false; }
    } else {
        mean1 = (k >= u1l && k <= nl1n - 1);
        mean2 = (k >= u2l && k <= nl2n - 1);
    }
    const int ua0 = ri[ER_UA0], ua1 = ri[ER_UA1], ua2 = ri[ER_UA2];
    const int da0 = ri[ER_DA0], da1 = ri[ER_DA1], da2 = ri[ER_DA2];
    const size_t o1 = (size_t)n1 * L + k - 1, o2 = (size_t)n2 * L + k - 1;
    const size_t ei = (size_t)e * L + k - 1;
    const size_t kk = (size_t)k - 1;
    const int trb = blockIdx.y * tpb, tre = min(t.ntr, trb + tpb);
    for (int tr = trb; tr < tre; ++tr) {
        const double *T = t.T + t.sN * tr;
        const double *Told = t.Told + t.sOld * tr;
        const int abo = t.abo[tr], lim = t.lim[tr], hor = t.hor[tr];
        double flo = 0.0;
        const double T1 = T[o1], T2 = T[o2];
        if (lim == LIM_FCT) {
            if (s1 >= 0) flo = upw1_h(T1, T2, vf1, flo);
            if (s2 >= 0) flo = upw1_h(T1, T2, vf2, flo);
            t.lo_h[t.sE * tr + ei] = flo;
        }
        double A1, A2;
        if (abo == 2) {
            A1 = -(0.5 + FGPU_EPS_AB) * Told[o1] + (1.5 + FGPU_EPS_AB) * T1;
            A2 = -(0.5 + FGPU_EPS_AB) * Told[o2] + (1.5 + FGPU_EPS_AB) * T2;
        } else {
            A1 = t_ab(abo, T, Told, o1);
            A2 = t_ab(abo, T, Told, o2);
        }
        double f = flo;
        if (hor != HOR_UPW1) {   // MFCT, MUSCL
            const double no = t.ph[tr];
            double g[4] = {0.0, 0.0, 0.0, 0.0};
            if (shared) {
                const double u0 = T[(size_t)ua0 * L + kk], u1_ = T[(size_t)ua1 * L + kk], u2_ = T[(size_t)ua2 * L + kk];
                const double w0 = T[(size_t)da0 * L + kk], w1 = T[(size_t)da1 * L + kk], w2 = T[(size_t)da2 * L + kk];
                g[0] = rd[ER_GU0] * u0 + rd[ER_GU1] * u1_ + rd[ER_GU2] * u2_;
                g[2] = rd[ER_GU3] * u0 + rd[ER_GU4] * u1_ + rd[ER_GU5] * u2_;
                g[1] = rd[ER_GD0] * w0 + rd[ER_GD1] * w1 + rd[ER_GD2] * w2;
                g[3] = rd[ER_GD3] * w0 + rd[ER_GD4] * w1 + rd[ER_GD5] * w2;
            } else {
                const double *gm = t.gmean + t.sG * tr;
                if (mean1) { g[0] = gm[2 * o1]; g[2] = gm[2 * o1 + 1]; }
                if (mean2) { g[1] = gm[2 * o2]; g[3] = gm[2 * o2 + 1]; }
            }
            double Tmean2, Tmean1;
            if (hor == HOR_MUSCL) {   // reconstruction only down to nboundary_lay (oce_adv_tra_hor.F90:356-367)
                const double c1 = k <= ri[ER_NBL1] ? 1.0 : 0.0, c2 = k <= ri[ER_NBL2] ? 1.0 : 0.0;
                Tmean2 = A2 - (2.0 * (A2 - A1) + edx * a * g[1] + edy * FGPU_R_EARTH * g[3]) / 6.0 * c2;
                Tmean1 = A1 + (2.0 * (A2 - A1) + edx * a * g[0] + edy * FGPU_R_EARTH * g[2]) / 6.0 * c1;
            } else {
                Tmean2 = A2 - (2.0 * (A2 - A1) + edx * a * g[1] + edy * FGPU_R_EARTH * g[3]) / 6.0;
                Tmean1 = A1 + (2.0 * (A2 - A1) + edx * a * g[0] + edy * FGPU_R_EARTH * g[2]) / 6.0;
            }
            for (int sg = 0; sg < 2; ++sg) {
                if ((sg == 0 ? s1 : s2) < 0) continue;
                const double vflux = sg == 0 ? vf1 : vf2;
                const double cHO = (vflux + fabs(vflux)) * Tmean1 + (vflux - fabs(vflux)) * Tmean2;
                f = -0.5 * (1.0 - no) * cHO - vflux * no * 0.5 * (Tmean1 + Tmean2) - f;
            }
        } else {
            if (s1 >= 0) f = upw1_h(A1, A2, vf1, f);
            if (s2 >= 0) f = upw1_h(A1, A2, vf2, f);
        }
        t.a_h[t.sE * tr + ei] = f;
    }
}

__device__ __forceinline__ double2 ld2(const double *p) { return *(const double2 *)p; }

// store levels k (x) and k + 1 (y) of a row, those with their flag set: both
// with one 16-byte store, one of them with an 8-byte store
__device__ __forceinline__ void st2(double *p, double x, double y, bool a0, bool a1)
{
    if (a0 && a1) {
        *(double2 *)p = make_double2(x, y);
    } else if (a0) {
        p[0] = x;
    } else if (a1) {
        p[1] = y;
    }
}

// ------------------------------------------------------------------------------
// K_A with two adjacent levels per lane: lane j takes levels 2j+1 and 2j+2, so
// one wave covers an edge of up to 128 levels, and each per-level row is read
// and written with 16-byte accesses (an even number of layers keeps every row
// 16-byte aligned). A 16-byte load per lane costs the vector memory pipeline
// what an 8-byte one does, so this halves its work against kA_edge<128>. Each
// level's arithmetic is kA_edge's, in the same order.
// ------------------------------------------------------------------------------
__global__ void __launch_bounds__(64) kA2_edge(DevMesh m, DevDyn d, DevTracers t, int tpb)
{
    const int je = xcd_column(m.E2);
    if (je >= m.E2) return;
    const int e = m.edge_ord[je];
    const int kb = 2 * (int)threadIdx.x + 1;   // this lane's levels: kb, kb + 1
    const int L = m.L;
    const int *ri = m.erec_i + (size_t)e * ERI;
    const double *rd = m.erec_d + (size_t)e * ERD;
    const int nl1 = ri[ER_NL1], nu1 = ri[ER_NU1], nl2 = ri[ER_NL2], nu2 = ri[ER_NU2];
    const int nl12 = min(nl1, nl2), nu12 = max(nu1, nu2);
    bool act[2], need2[2];
    int s1[2], s2[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int k = kb + s;
        const bool i0 = (k >= nu1 && k <= nu12 - 1);
        const bool i1 = (nu2 > 0 && k >= nu2 && k <= nu12 - 1);
        const bool i2 = (k >= nu12 && k <= nl12);
        const bool i3 = (k >= nl12 + 1 && k <= nl1);
        const bool i4 = (k >= nl12 + 1 && k <= nl2);
        act[s] = i0 || i1 || i2 || i3 || i4;
        need2[s] = i1 || i2 || i4;
        // at most two segments meet at one level: A then B, or C, or D then E
        s1[s] = i0 ? 0 : (i2 ? 2 : (i3 ? 3 : -1));
        s2[s] = i1 ? 1 : (i4 ? 4 : -1);
    }
    if (!act[0] && !act[1]) return;
    const int n1 = ri[ER_N1], n2 = ri[ER_N2], el1 = ri[ER_EL1], el2 = ri[ER_EL2];
    const double dX1 = rd[ER_DX1], dY1 = rd[ER_DY1], dX2 = rd[ER_DX2], dY2 = rd[ER_DY2];
    const double a = rd[ER_A], edx = rd[ER_EDX], edy = rd[ER_EDY];
    double vf1[2], vf2[2];
    {
        const size_t q1 = (size_t)el1 * L + kb - 1;
        const double2 uva = ld2(d.uv + 2 * q1), uvb = ld2(d.uv + 2 * q1 + 2), hh1 = ld2(m.helem + q1);
        const double u1[2] = {uva.x, uvb.x}, v1[2] = {uva.y, uvb.y}, h1[2] = {hh1.x, hh1.y};
        double u2[2] = {0.0, 0.0}, v2[2] = {0.0, 0.0}, h2[2] = {0.0, 0.0};
        if (el2 >= 0 && ((act[0] && need2[0]) || (act[1] && need2[1]))) {
            const size_t q2 = (size_t)el2 * L + kb - 1;
            const double2 wa = ld2(d.uv + 2 * q2), wb = ld2(d.uv + 2 * q2 + 2), hh2 = ld2(m.helem + q2);
            if (need2[0]) { u2[0] = wa.x; v2[0] = wa.y; h2[0] = hh2.x; }
            if (need2[1]) { u2[1] = wb.x; v2[1] = wb.y; h2[1] = hh2.y; }
        }
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            vf1[s] = s1[s] >= 0 ? seg_vflux(s1[s], u1[s], v1[s], u2[s], v2[s], h1[s], h2[s], dX1, dY1, dX2, dY2) : 0.0;
            vf2[s] = s2[s] >= 0 ? seg_vflux(s2[s], u1[s], v1[s], u2[s], v2[s], h1[s], h2[s], dX1, dY1, dX2, dY2) : 0.0;
        }
    }
    // gradient rule of fill_up_dn_grad per level (oce_muscl_adv.F90:375-520)
    bool shared[2], mean1[2], mean2[2];
    const int u1l = ri[ER_U1], u2l = ri[ER_U2], nl1n = ri[ER_NL1N], nl2n = ri[ER_NL2N];
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int k = kb + s;
        shared[s] = false;
        if (ri[ER_RULE]) {
            const int nzmin = ri[ER_NZMIN], nzmax = ri[ER_NZMAX];
            mean1[s] = (k >= u1l && k <= nzmin - 1);
            mean2[s] = (k >= u2l && k <= nzmin - 1);
            if (k >= nzmin && k <= nzmax - 1) { shared[s] = true; mean1[s] = false; mean2[s] = false; }
            if (k >= nzmax && k <= nl1n - 1) { mean1[s] = true; shared[s] = false; }
            if (k >= nzmax && k <= nl2n - 1) { mean2[s] = true; shared[s] = false; }
        } else {
            mean1[s] = (k >= u1l && k <= nl1n - 1);
            mean2[s] = (k >= u2l && k <= nl2n - 1);
        }
    }
    const bool any_sh = (act[0] && shared[0]) || (act[1] && shared[1]);
    const bool any_m1 = (act[0] && !shared[0] && mean1[0]) || (act[1] && !shared[1] && mean1[1]);
    const bool any_m2 = (act[0] && !shared[0] && mean2[0]) || (act[1] && !shared[1] && mean2[1]);
    const int ua0 = ri[ER_UA0], ua1 = ri[ER_UA1], ua2 = ri[ER_UA2];
    const int da0 = ri[ER_DA0], da1 = ri[ER_DA1], da2 = ri[ER_DA2];
    const size_t o1 = (size_t)n1 * L + kb - 1, o2 = (size_t)n2 * L + kb - 1;
    const size_t ei = (size_t)e * L + kb - 1;
    const size_t kk = (size_t)kb - 1;
    const int trb = blockIdx.y * tpb, tre = min(t.ntr, trb + tpb);
    for (int tr = trb; tr < tre; ++tr) {
        const double *T = t.T + t.sN * tr;
        const double *Told = t.Told + t.sOld * tr;
        const int abo = t.abo[tr], lim = t.lim[tr], hor = t.hor[tr];
        const double2 t1 = ld2(T + o1), t2 = ld2(T + o2);
        const double T1[2] = {t1.x, t1.y}, T2[2] = {t2.x, t2.y};
        double flo[2] = {0.0, 0.0};
        if (lim == LIM_FCT) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (s1[s] >= 0) flo[s] = upw1_h(T1[s], T2[s], vf1[s], flo[s]);
                if (s2[s] >= 0) flo[s] = upw1_h(T1[s], T2[s], vf2[s], flo[s]);
            }
            st2(t.lo_h + t.sE * tr + ei, flo[0], flo[1], act[0], act[1]);
        }
        double A1[2], A2[2];   // T^AB (oce_tracer_mod.F90:44-59)
        if (abo == 2) {
            const double2 p1 = ld2(Told + o1), p2 = ld2(Told + o2);
            const double O1[2] = {p1.x, p1.y}, O2[2] = {p2.x, p2.y};
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                A1[s] = -(0.5 + FGPU_EPS_AB) * O1[s] + (1.5 + FGPU_EPS_AB) * T1[s];
                A2[s] = -(0.5 + FGPU_EPS_AB) * O2[s] + (1.5 + FGPU_EPS_AB) * T2[s];
            }
        } else {
            const double2 p1a = ld2(Told + 2 * o1), p1b = ld2(Told + 2 * o1 + 2);
            const double2 p2a = ld2(Told + 2 * o2), p2b = ld2(Told + 2 * o2 + 2);
            A1[0] = (5.0 * p1a.y - 16.0 * p1a.x + 23.0 * T1[0]) / 12.0;
            A1[1] = (5.0 * p1b.y - 16.0 * p1b.x + 23.0 * T1[1]) / 12.0;
            A2[0] = (5.0 * p2a.y - 16.0 * p2a.x + 23.0 * T2[0]) / 12.0;
            A2[1] = (5.0 * p2b.y - 16.0 * p2b.x + 23.0 * T2[1]) / 12.0;
        }
        double f[2] = {flo[0], flo[1]};
        if (hor != HOR_UPW1) {   // MFCT, MUSCL
            const double no = t.ph[tr];
            double2 ux0, ux1, ux2, wx0, wx1, wx2;
            if (any_sh) {
                ux0 = ld2(T + (size_t)ua0 * L + kk);
                ux1 = ld2(T + (size_t)ua1 * L + kk);
                ux2 = ld2(T + (size_t)ua2 * L + kk);
                wx0 = ld2(T + (size_t)da0 * L + kk);
                wx1 = ld2(T + (size_t)da1 * L + kk);
                wx2 = ld2(T + (size_t)da2 * L + kk);
            }
            const double *gm = t.gmean + t.sG * tr;
            double2 m1a, m1b, m2a, m2b;
            if (any_m1) { m1a = ld2(gm + 2 * o1); m1b = ld2(gm + 2 * o1 + 2); }
            if (any_m2) { m2a = ld2(gm + 2 * o2); m2b = ld2(gm + 2 * o2 + 2); }
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                double g[4] = {0.0, 0.0, 0.0, 0.0};
                if (shared[s]) {
                    const double u0 = s ? ux0.y : ux0.x, u1_ = s ? ux1.y : ux1.x, u2_ = s ? ux2.y : ux2.x;
                    const double w0 = s ? wx0.y : wx0.x, w1 = s ? wx1.y : wx1.x, w2 = s ? wx2.y : wx2.x;
                    g[0] = rd[ER_GU0] * u0 + rd[ER_GU1] * u1_ + rd[ER_GU2] * u2_;
                    g[2] = rd[ER_GU3] * u0 + rd[ER_GU4] * u1_ + rd[ER_GU5] * u2_;
                    g[1] = rd[ER_GD0] * w0 + rd[ER_GD1] * w1 + rd[ER_GD2] * w2;
                    g[3] = rd[ER_GD3] * w0 + rd[ER_GD4] * w1 + rd[ER_GD5] * w2;
                } else {
                    if (mean1[s]) { const double2 q = s ? m1b : m1a; g[0] = q.x; g[2] = q.y; }
                    if (mean2[s]) { const double2 q = s ? m2b : m2a; g[1] = q.x; g[3] = q.y; }
                }
                double Tmean2, Tmean1;
                if (hor == HOR_MUSCL) {   // reconstruction only down to nboundary_lay (oce_adv_tra_hor.F90:356-367)
                    const int k = kb + s;
                    const double c1 = k <= ri[ER_NBL1] ? 1.0 : 0.0, c2 = k <= ri[ER_NBL2] ? 1.0 : 0.0;
                    Tmean2 = A2[s] - (2.0 * (A2[s] - A1[s]) + edx * a * g[1] + edy * FGPU_R_EARTH * g[3]) / 6.0 * c2;
                    Tmean1 = A1[s] + (2.0 * (A2[s] - A1[s]) + edx * a * g[0] + edy * FGPU_R_EARTH * g[2]) / 6.0 * c1;
                } else {
                    Tmean2 = A2[s] - (2.0 * (A2[s] - A1[s]) + edx * a * g[1] + edy * FGPU_R_EARTH * g[3]) / 6.0;
                    Tmean1 = A1[s] + (2.0 * (A2[s] - A1[s]) + edx * a * g[0] + edy * FGPU_R_EARTH * g[2]) / 6.0;
                }
                for (int sg = 0; sg < 2; ++sg) {
                    if ((sg == 0 ? s1[s] : s2[s]) < 0) continue;
                    const double vflux = sg == 0 ? vf1[s] : vf2[s];
                    const double cHO = (vflux + fabs(vflux)) * Tmean1 + (vflux - fabs(vflux)) * Tmean2;
                    f[s] = -0.5 * (1.0 - no) * cHO - vflux * no * 0.5 * (Tmean1 + Tmean2) - f[s];
                }
            }
        } else {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                if (s1[s] >= 0) f[s] = upw1_h(A1[s], A2[s], vf1[s], f[s]);
                if (s2[s] >= 0) f[s] = upw1_h(A1[s], A2[s], vf2[s], f[s]);
            }
        }
        st2(t.a_h + t.sE * tr + ei, f[0], f[1], act[0], act[1]);
    }
}

// ------------------------------------------------------------------------------
// column helpers on LDS copies (index = level - 1)
// ------------------------------------------------------------------------------
__device__ __forceinline__ double col_upw1(int j, int nzmin, int nzmax, const double *T, const double *W,
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

__device__ __forceinline__ double col_qr4c(int j, int nzmin, int nzmax, const double *T, const double *W,
                                           const double *ar, const double *Zn, const double *zb, double num_ord,
                                           double f)
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

// ------------------------------------------------------------------------------
// K_B: LO solution and antidiffusive vertical flux per node column
// ------------------------------------------------------------------------------
// XV: the CDIFF and PPM vertical schemes are compiled in (launched only when a
// tracer uses one; their registers and LDS would slow the QR4C/UPW1 version).
// WS: use_wsplit, the LO solution of FCT tracers takes the implicit upwind
// advection with w_i (adv_tra_vert_impl, oce_adv_tra_ver.F90:90-240) and the
// HO scheme is handed the upwind flux with the full W.
template <int LB, bool XV, bool WS>
__global__ void __launch_bounds__(LB) kB_node(DevMesh m, DevDyn d, DevTracers t, double dt, int tpb)
{
    const int j = xcd_column(m.N2);
    if (j >= m.N2) return;
    const int n = m.node_ord[j];
    __shared__ double sT[LB + 2], sA[LB + 2], sW[LB + 2], sWe[LB + 2], sAr[LB + 2], sZ[LB + 2], sZb[LB + 2];
    __shared__ double sLo[LB + 2], sHn[XV ? LB + 2 : 1], sHo[XV ? LB + 2 : 1];
    constexpr int NW = WS ? LB + 2 : 1;
    __shared__ double sIa[NW], sIb[NW], sIc[NW], sIcp[NW], sIf[NW], sIr[NW];
    const int tid = threadIdx.x, k = tid + 1;
    const int L = m.L, NL = m.NL;
    const int nzmin = m.ulevn[n], nzmax = m.nlevn[n];
    const bool act = (k >= nzmin && k <= nzmax - 1);
    const size_t ci = (size_t)n * L + k - 1;
    const NodeEdges E = load_node_edges(m, n);
    if (k <= NL) {
        sW[tid] = d.w[(size_t)n * NL + tid];
        sWe[tid] = d.w_e[(size_t)n * NL + tid];
        sAr[tid] = m.area[(size_t)n * NL + tid];
        sZb[tid] = m.zbar3d[(size_t)n * NL + tid];
    }
    if (k <= L) {
        sZ[tid] = m.Z3d[ci];
        if (XV) {   // PPM: thicknesses of the neighbouring layers
            sHn[tid] = m.hnode_new[ci];
            sHo[tid] = m.hnode[ci];
        }
    }
    double hn = 1.0, hh = 0.0, av_ = 1.0;
    if (act) {
        hn = m.hnode_new[ci];
        hh = m.hnode[ci];
        av_ = m.areasvol[(size_t)n * NL + k - 1];
    }
    if (WS) __syncthreads();   // sAr
    if (WS && act) {   // tridiagonal rows of the implicit upwind step (tracer independent)
        const double *wi = d.w_i + (size_t)n * NL;
        const double zinv = 1.0 * dt;
        double v_adv = zinv * sAr[k - 1] / av_;
        double a, b, c;
        if (k == nzmin) {
            a = 0.0;
            b = hn + wi[k - 1] * v_adv;
        } else {
            a = dmn(0.0, wi[k - 1]) * v_adv;
            b = hn + dmx(0.0, wi[k - 1]) * v_adv;
        }
        if (k <= nzmax - 2) {
            v_adv = zinv * sAr[k] / av_;
            b = b - dmn(0.0, wi[k]) * v_adv;
            c = -dmx(0.0, wi[k]) * v_adv;
        } else {
            c = 0.0;
        }
        sIa[tid] = a;
        sIb[tid] = b;
        sIc[tid] = c;
    }
    const int trb = blockIdx.y * tpb, tre = min(t.ntr, trb + tpb);
    for (int tr = trb; tr < tre; ++tr) {
        const double *T = t.T + t.sN * tr;
        const double *Told = t.Told + t.sOld * tr;
        const int abo = t.abo[tr], lim = t.lim[tr], ver = t.ver[tr];
        const double pv = t.pv[tr];
        __syncthreads();
        if (k <= L) {
            sT[tid] = T[ci];
            sA[tid] = t_ab(abo, T, Told, ci);
        }
        __syncthreads();
        double *av = t.a_v + t.sNI * tr + (size_t)n * NL;
        const double *W = (lim == LIM_FCT) ? sW : sWe;
        double lk = 0.0;
        if (lim == LIM_FCT) {
            if (k >= nzmin && k <= nzmax) lk = col_upw1(k, nzmin, nzmax, sT, sWe, sAr, 0.0);
            sLo[tid] = lk;
        }
        __syncthreads();
        if (k >= nzmin && k <= nzmax) {
            // LO flux subtracted from the HO one: upwind with the full W under the split
            const double fk = (WS && lim == LIM_FCT) ? col_upw1(k, nzmin, nzmax, sT, sW, sAr, 0.0) : lk;
            double ho;
            if (ver == VER_QR4C) ho = col_qr4c(k, nzmin, nzmax, sA, W, sAr, sZ, sZb, pv, fk);
            else if (XV && ver == VER_CDIFF) ho = ver_cdiff(k, nzmin, nzmax, sA, W, sAr, fk);
            else if (XV && ver == VER_PPM) ho = ver_ppm(k, nzmin, nzmax, sA, W, sAr, sHn, sHo, dt, fk);
            else ho = col_upw1(k, nzmin, nzmax, sA, W, sAr, fk);
            av[k - 1] = ho;
        }
        if (lim == LIM_FCT) {
            const double *lo_h = t.lo_h + t.sE * tr;
            double acc = 0.0;
            // edges in batches of NG: the batch's loads are issued together
            // (clamped to valid addresses), then added in ascending edge order
            const int kk = min(k, L) - 1;
            for (int j0 = 0; j0 < E.deg; j0 += NG) {
                double fl_[NG];
                int q_[NG];
                bool ok_[NG];
#pragma unroll
                for (int u = 0; u < NG; ++u) {
                    const bool real = j0 + u < E.deg;
                    const int j = real ? j0 + u : 0;
                    const int q = rl(E.q, j), a0 = rl(E.nu12, j), a1 = rl(E.nl12, j);
                    ok_[u] = real && act && k >= a0 && k <= a1;
                    q_[u] = q;
                    fl_[u] = lo_h[(size_t)(q >> 1) * L + kk];
                }
#pragma unroll
                for (int u = 0; u < NG; ++u)
                    if (ok_[u]) acc = (q_[u] & 1) ? acc - fl_[u] : acc + fl_[u];
            }
            if (!WS) {
                if (act) {
                    const double lk1 = sLo[tid + 1];
                    t.fct_lo[t.sN * tr + ci] = (sT[tid] * hh + (acc + (lk - lk1)) * dt / av_) / hn;
                }
            } else {
                if (act) {
                    const double lk1 = sLo[tid + 1];
                    sIf[tid] = (sT[tid] * hh + (acc + (lk - lk1)) * dt / av_) / hn;
                }
                __syncthreads();
                if (act) {   // right-hand side of the implicit step
                    double r;
                    if (k == nzmin) r = -(sIb[tid] - hn) * sIf[tid] - sIc[tid] * sIf[tid + 1];
                    else if (k <= nzmax - 2) r = -sIa[tid] * sIf[tid - 1] - (sIb[tid] - hn) * sIf[tid] - sIc[tid] * sIf[tid + 1];
                    else r = -sIa[tid] * sIf[tid - 1] - (sIb[tid] - hn) * sIf[tid];
                    sIr[tid] = r;
                }
                __syncthreads();
                if (tid == 0) {   // Thomas sweeps of the column, one lane
                    double cp = sIc[nzmin - 1] / sIb[nzmin - 1];
                    double tp = sIr[nzmin - 1] / sIb[nzmin - 1];
                    sIcp[nzmin - 1] = cp;
                    sIr[nzmin - 1] = tp;
                    for (int nz = nzmin + 1; nz <= nzmax - 1; ++nz) {
                        const double an = sIa[nz - 1];
                        const double mm = sIb[nz - 1] - cp * an;
                        cp = sIc[nz - 1] / mm;
                        tp = (sIr[nz - 1] - tp * an) / mm;
                        sIcp[nz - 1] = cp;
                        sIr[nz - 1] = tp;
                    }
                    double x = tp;
                    for (int nz = nzmax - 2; nz >= nzmin; --nz) {
                        x = sIr[nz - 1] - sIcp[nz - 1] * x;
                        sIr[nz - 1] = x;
                    }
                }
                __syncthreads();
                if (act) t.fct_lo[t.sN * tr + ci] = sIf[tid] + sIr[tid];
            }
        }
    }
}

// ------------------------------------------------------------------------------
// K_B with two adjacent levels per lane: lane j takes levels and interfaces
// 2j+1 and 2j+2, so one wave covers a column of up to 128 levels, and the
// per-level rows (T, valuesold, Z, the thicknesses, the edges' LO fluxes and
// fct_LO) move with 16-byte accesses (an even number of layers keeps those
// rows 16-byte aligned). The interface rows (w, area, zbar, A_v) keep 8-byte
// accesses: with an odd interface count every other column's row starts
// 8 bytes off. Each level's arithmetic is kB_node's, in the same order. For
// the step without the W split and without CDIFF or PPM tracers.
// ------------------------------------------------------------------------------
__global__ void __launch_bounds__(64) kB2_node(DevMesh m, DevDyn d, DevTracers t, double dt, int tpb)
{
    const int jc = xcd_column(m.N2);
    if (jc >= m.N2) return;
    const int n = m.node_ord[jc];
    __shared__ double sT[130], sA[130], sW[130], sWe[130], sAr[130], sZ[130], sZb[130], sLo[130];
    const int kb = 2 * (int)threadIdx.x + 1;   // this lane's levels and interfaces: kb, kb + 1
    const int L = m.L, NL = m.NL;
    const int nzmin = m.ulevn[n], nzmax = m.nlevn[n];
    bool act[2];
#pragma unroll
    for (int s = 0; s < 2; ++s) act[s] = kb + s >= nzmin && kb + s <= nzmax - 1;
    const size_t ci = (size_t)n * L + kb - 1;   // even
    const NodeEdges E = load_node_edges(m, n);
#pragma unroll
    for (int s = 0; s < 2; ++s) {
        const int k = kb + s;
        if (k <= NL) {
            const size_t ii = (size_t)n * NL + k - 1;
            sW[k - 1] = d.w[ii];
            sWe[k - 1] = d.w_e[ii];
            sAr[k - 1] = m.area[ii];
            sZb[k - 1] = m.zbar3d[ii];
        }
    }
    if (kb <= L) {
        const double2 z = ld2(m.Z3d + ci);
        sZ[kb - 1] = z.x;
        sZ[kb] = z.y;
    }
    double hn[2] = {1.0, 1.0}, hh[2] = {0.0, 0.0}, av_[2] = {1.0, 1.0};
    if (act[0] || act[1]) {
        const double2 a = ld2(m.hnode_new + ci), b = ld2(m.hnode + ci);
        if (act[0]) { hn[0] = a.x; hh[0] = b.x; av_[0] = m.areasvol[(size_t)n * NL + kb - 1]; }
        if (act[1]) { hn[1] = a.y; hh[1] = b.y; av_[1] = m.areasvol[(size_t)n * NL + kb]; }
    }
    const size_t kk = (size_t)min(kb, L - 1) - 1;   // even; the row's last pair for lanes past L
    const int trb = blockIdx.y * tpb, tre = min(t.ntr, trb + tpb);
    for (int tr = trb; tr < tre; ++tr) {
        const double *T = t.T + t.sN * tr;
        const double *Told = t.Told + t.sOld * tr;
        const int abo = t.abo[tr], lim = t.lim[tr], ver = t.ver[tr];
        const double pv = t.pv[tr];
        __syncthreads();
        if (kb <= L) {   // T and T^AB (oce_tracer_mod.F90:44-59) of both levels
            const double2 tv = ld2(T + ci);
            sT[kb - 1] = tv.x;
            sT[kb] = tv.y;
            if (abo == 2) {
                const double2 o = ld2(Told + ci);
                sA[kb - 1] = -(0.5 + FGPU_EPS_AB) * o.x + (1.5 + FGPU_EPS_AB) * tv.x;
                sA[kb] = -(0.5 + FGPU_EPS_AB) * o.y + (1.5 + FGPU_EPS_AB) * tv.y;
            } else {
                const double2 oa = ld2(Told + 2 * ci), ob = ld2(Told + 2 * ci + 2);
                sA[kb - 1] = (5.0 * oa.y - 16.0 * oa.x + 23.0 * tv.x) / 12.0;
                sA[kb] = (5.0 * ob.y - 16.0 * ob.x + 23.0 * tv.y) / 12.0;
            }
        }
        __syncthreads();
        double *av = t.a_v + t.sNI * tr + (size_t)n * NL;
        const double *W = (lim == LIM_FCT) ? sW : sWe;
        double lk[2] = {0.0, 0.0};
        if (lim == LIM_FCT) {
#pragma unroll
            for (int s = 0; s < 2; ++s) {
                const int k = kb + s;
                if (k >= nzmin && k <= nzmax) lk[s] = col_upw1(k, nzmin, nzmax, sT, sWe, sAr, 0.0);
                sLo[k - 1] = lk[s];
            }
        }
        __syncthreads();
#pragma unroll
        for (int s = 0; s < 2; ++s) {
            const int k = kb + s;
            if (k >= nzmin && k <= nzmax) {
                double ho;
                if (ver == VER_QR4C) ho = col_qr4c(k, nzmin, nzmax, sA, W, sAr, sZ, sZb, pv, lk[s]);
                else ho = col_upw1(k, nzmin, nzmax, sA, W, sAr, lk[s]);
                av[k - 1] = ho;
            }
        }
        if (lim == LIM_FCT) {
            const double *lo_h = t.lo_h + t.sE * tr;
            double acc[2] = {0.0, 0.0};
            // edges in batches of NG: the batch's loads are issued together,
            // then added in ascending edge order
            for (int j0 = 0; j0 < E.deg; j0 += NG) {
                double2 fl_[NG];
                int q_[NG], a0_[NG], a1_[NG];
                bool real_[NG];
#pragma unroll
                for (int u = 0; u < NG; ++u) {
                    real_[u] = j0 + u < E.deg;
                    const int j = real_[u] ? j0 + u : 0;
                    q_[u] = rl(E.q, j);
                    a0_[u] = rl(E.nu12, j);
                    a1_[u] = rl(E.nl12, j);
                    fl_[u] = ld2(lo_h + (size_t)(q_[u] >> 1) * L + kk);
                }
#pragma unroll
                for (int u = 0; u < NG; ++u) {
#pragma unroll
                    for (int s = 0; s < 2; ++s) {
                        const int k = kb + s;
                        if (real_[u] && act[s] && k >= a0_[u] && k <= a1_[u]) {
                            const double fl = s ? fl_[u].y : fl_[u].x;
                            acc[s] = (q_[u] & 1) ? acc[s] - fl : acc[s] + fl;
                        }
                    }
                }
            }
            double o[2] = {0.0, 0.0};
#pragma unroll
            for (int s = 0; s < 2; ++s)
                if (act[s]) o[s] = (sT[kb - 1 + s] * hh[s] + (acc[s] + (lk[s] - sLo[kb + s])) * dt / av_[s]) / hn[s];
            st2(t.fct_lo + t.sN * tr + ci, o[0], o[1], act[0], act[1]);
        }
    }
}

// ------------------------------------------------------------------------------
// K_C: FCT bounds, P+/P-, R+/R- per node column (oce_adv_tra_fct.F90:119-405)
// The element bound (max over the 3 nodes of max(LO, T^n)), maximised over the
// elements around the node, equals the maximum over the node and those edge
// neighbours m for which an element adjacent to edge (n, m) has the level,
// together with the padding -1e3 (max) / +1e3 (min) of every element around
// the node that lacks the level (oce_adv_tra_fct.F90:148-178).
// ------------------------------------------------------------------------------
template <int LB>
__global__ void __launch_bounds__(LB) kC_node(DevMesh m, DevTracers t, double dt, int tpb)
{
    const int j = xcd_column(m.N2);
    if (j >= m.N2) return;
    const int n = m.node_ord[j];
    __shared__ double sMx[LB + 2], sMn[LB + 2];
    const int tid = threadIdx.x, k = tid + 1;
    const int L = m.L, NL = m.NL;
    const int nzmin = m.ulevn[n], nzmax = m.nlevn[n];
    const bool act = (k >= nzmin && k <= nzmax - 1);
    const size_t ci = (size_t)n * L + k - 1;
    const NodeEdges E = load_node_edges(m, n);
    double hn = 1.0, av_ = 1.0;
    if (act) {
        hn = m.hnode_new[ci];
        av_ = m.areasvol[(size_t)n * NL + k - 1];
    }
    const int2 pr = *(const int2 *)(m.fct_pad + 2 * (size_t)n);
    const bool padded = k < pr.x || k >= pr.y;   // an element around n lacks level k
    const int trb = blockIdx.y * tpb, tre = min(t.ntr, trb + tpb);
    for (int tr = trb; tr < tre; ++tr) {
        if (t.lim[tr] != LIM_FCT) continue;   // uniform over the block
        const double *T = t.T + t.sN * tr;
        const double *LO = t.fct_lo + t.sN * tr;
        const double *ah = t.a_h + t.sE * tr;
        double vmax = -1e3, vmin = 1e3, lo = 0.0, p = 0.0, q = 0.0;
        if (act) {
            lo = LO[ci];
            const double tk = T[ci];
            vmax = dmx(lo, tk);
            vmin = dmn(lo, tk);
            if (padded) {
                vmax = dmx(vmax, -1e3);
                vmin = dmn(vmin, 1e3);
            }
            const double *av = t.a_v + t.sNI * tr + (size_t)n * NL;
            p = p + (dmx(0.0, av[k - 1]) + dmx(0.0, -av[k]));
            q = q + (dmn(0.0, av[k - 1]) + dmn(0.0, -av[k]));
        }
        const int kk = min(k, L) - 1;
        for (int j0 = 0; j0 < E.deg; j0 += NG) {
            double lm_[NG], tm_[NG], fl_[NG];
            int qq_[NG];
            bool has_[NG], in_[NG];
#pragma unroll
            for (int u = 0; u < NG; ++u) {
                const bool real = j0 + u < E.deg;
                const int j = real ? j0 + u : 0;
                const int qq = rl(E.q, j), mo = rl(E.other, j);
                const int a0 = rl(E.nu12, j), a1 = rl(E.nl12, j);
                const int b_nu1 = rl(E.nu1, j), b_nl1 = rl(E.nl1, j), b_nu2 = rl(E.nu2, j), b_nl2 = rl(E.nl2, j);
                has_[u] = real && act && ((k >= b_nu1 && k <= b_nl1) || (b_nu2 > 0 && k >= b_nu2 && k <= b_nl2));
                in_[u] = real && act && k >= a0 && k <= a1;
                qq_[u] = qq;
                const size_t mi = (size_t)mo * L + kk;
                lm_[u] = LO[mi];
                tm_[u] = T[mi];
                fl_[u] = ah[(size_t)(qq >> 1) * L + kk];
            }
#pragma unroll
            for (int u = 0; u < NG; ++u) {
                if (has_[u]) {
                    vmax = dmx(vmax, dmx(lm_[u], tm_[u]));
                    vmin = dmn(vmin, dmn(lm_[u], tm_[u]));
                }
                if (in_[u]) {
                    const double fl = fl_[u];
                    if (qq_[u] & 1) {
                        p = p + dmx(0.0, -fl);
                        q = q + dmn(0.0, -fl);
                    } else {
                        p = p + dmx(0.0, fl);
                        q = q + dmn(0.0, fl);
                    }
                }
            }
        }
        __syncthreads();
        sMx[tid] = vmax;
        sMn[tid] = vmin;
        __syncthreads();
        if (act) {
            double bmax, bmin;
            if (k == nzmax - 1 || k == nzmin) {
                bmax = vmax - lo;
                bmin = vmin - lo;
            } else {
                bmax = dmx(dmx(sMx[tid - 1], vmax), sMx[tid + 1]) - lo;
                bmin = dmn(dmn(sMn[tid - 1], vmin), sMn[tid + 1]) - lo;
            }
            double flux = p * dt / av_ / hn + 1e-16;
            t.rp[t.sN * tr + ci] = dmn(1.0, bmax / flux);
            flux = q * dt / av_ / hn - 1e-16;
            t.rm[t.sN * tr + ci] = dmn(1.0, bmin / flux);
        }
    }
}

// ------------------------------------------------------------------------------
// K_C over (tracer, level) items: one wave per node column and TPF tracers;
// lane j of chunk c takes item i = 64 c + j, tracer trb + i / nw and level
// nzmin + i % nw of the column's nw wet layers, so a wave instruction covers
// 64 wet (tracer, level) pairs instead of one tracer's column padded to the
// lane count. Same arithmetic per (tracer, level) as kC_node, neighbours added
// in the same order, so the results are bit-identical to it. NCH bounds the
// chunks: TPF x (levels of the deepest column) / 64, rounded up.
// ------------------------------------------------------------------------------
template <int TPF, int NCH>
__global__ void __launch_bounds__(64) kCf_node(DevMesh m, DevTracers t, double dt)
{
    constexpr int RS = 130;   // LDS row per tracer: levels 1..128 (index k - 1) and a guard
    __shared__ double sMx[TPF][RS], sMn[TPF][RS];
    const int jn = xcd_column(m.N2);
    if (jn >= m.N2) return;
    const int n = m.node_ord[jn];
    const int lane = threadIdx.x;
    const int L = m.L, NL = m.NL;
    const int nzmin = m.ulevn[n], nzmax = m.nlevn[n];
    const int nw = nzmax - nzmin;   // wet layers nzmin..nzmax-1
    if (nw < 1) return;
    const NodeEdges E = load_node_edges(m, n);
    const int2 pr = *(const int2 *)(m.fct_pad + 2 * (size_t)n);
    const int trb = blockIdx.y * TPF, tre = min(t.ntr, trb + TPF);
    const int items = (tre - trb) * nw;
    const float inw = 1.0f / (float)nw;
    double lo_[NCH], p_[NCH], q_[NCH];
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        lo_[c] = 0.0; p_[c] = 0.0; q_[c] = 0.0;
        const int i = 64 * c + lane;
        if (64 * c >= items) break;   // uniform
        const bool on = i < items;
        const int tri = on ? (int)(((float)i + 0.5f) * inw) : 0;
        const int tr = trb + tri, k = nzmin + (on ? i - tri * nw : 0);
        const bool act = on && t.lim[tr] == LIM_FCT;
        const size_t ci = (size_t)n * L + k - 1;
        const double *T = t.T + t.sN * tr;
        const double *LO = t.fct_lo + t.sN * tr;
        const double *ah = t.a_h + t.sE * tr;
        double vmax = -1e3, vmin = 1e3, lo = 0.0, p = 0.0, q = 0.0;
        if (act) {
            lo = LO[ci];
            const double tk = T[ci];
            vmax = dmx(lo, tk);
            vmin = dmn(lo, tk);
            if (k < pr.x || k >= pr.y) {   // an element around n lacks level k
                vmax = dmx(vmax, -1e3);
                vmin = dmn(vmin, 1e3);
            }
            const double *av = t.a_v + t.sNI * tr + (size_t)n * NL;
            p = p + (dmx(0.0, av[k - 1]) + dmx(0.0, -av[k]));
            q = q + (dmn(0.0, av[k - 1]) + dmn(0.0, -av[k]));
        }
        for (int j0 = 0; j0 < E.deg; j0 += NG) {
            double lm_[NG], tm_[NG], fl_[NG];
            int qq_[NG];
            bool has_[NG], in_[NG];
#pragma unroll
            for (int u = 0; u < NG; ++u) {
                const bool real = j0 + u < E.deg;
                const int j = real ? j0 + u : 0;
                const int qq = rl(E.q, j), mo = rl(E.other, j);
                const int a0 = rl(E.nu12, j), a1 = rl(E.nl12, j);
                const int b_nu1 = rl(E.nu1, j), b_nl1 = rl(E.nl1, j), b_nu2 = rl(E.nu2, j), b_nl2 = rl(E.nl2, j);
                has_[u] = real && act && ((k >= b_nu1 && k <= b_nl1) || (b_nu2 > 0 && k >= b_nu2 && k <= b_nl2));
                in_[u] = real && act && k >= a0 && k <= a1;
                qq_[u] = qq;
                const size_t mi = (size_t)mo * L + k - 1;
                lm_[u] = LO[mi];
                tm_[u] = T[mi];
                fl_[u] = ah[(size_t)(qq >> 1) * L + k - 1];
            }
#pragma unroll
            for (int u = 0; u < NG; ++u) {
                if (has_[u]) {
                    vmax = dmx(vmax, dmx(lm_[u], tm_[u]));
                    vmin = dmn(vmin, dmn(lm_[u], tm_[u]));
                }
                if (in_[u]) {
                    const double fl = fl_[u];
                    if (qq_[u] & 1) {
                        p = p + dmx(0.0, -fl);
                        q = q + dmn(0.0, -fl);
                    } else {
                        p = p + dmx(0.0, fl);
                        q = q + dmn(0.0, fl);
                    }
                }
            }
        }
        if (on) {
            sMx[tri][k - 1] = vmax;
            sMn[tri][k - 1] = vmin;
        }
        lo_[c] = lo;
        p_[c] = p;
        q_[c] = q;
    }
    wave_sync();
#pragma unroll
    for (int c = 0; c < NCH; ++c) {
        const int i = 64 * c + lane;
        if (64 * c >= items) break;
        const bool on = i < items;
        const int tri = on ? (int)(((float)i + 0.5f) * inw) : 0;
        const int tr = trb + tri, k = nzmin + (on ? i - tri * nw : 0);
        if (!(on && t.lim[tr] == LIM_FCT)) continue;
        const size_t ci = (size_t)n * L + k - 1;
        const double hn = m.hnode_new[ci], av_ = m.areasvol[(size_t)n * NL + k - 1];
        const double lo = lo_[c], vmax = sMx[tri][k - 1], vmin = sMn[tri][k - 1];
        double bmax, bmin;
        if (k == nzmax - 1 || k == nzmin) {
            bmax = vmax - lo;
            bmin = vmin - lo;
        } else {
            bmax = dmx(dmx(sMx[tri][k - 2], vmax), sMx[tri][k]) - lo;
            bmin = dmn(dmn(sMn[tri][k - 2], vmin), sMn[tri][k]) - lo;
        }
        double flux = p_[c] * dt / av_ / hn + 1e-16;
        t.rp[t.sN * tr + ci] = dmn(1.0, bmax / flux);
        flux = q_[c] * dt / av_ / hn - 1e-16;
        t.rm[t.sN * tr + ci] = dmn(1.0, bmin / flux);
    }
}

// ------------------------------------------------------------------------------
// K_D: limited fluxes, flux2dtracer, T*, implicit vertical diffusion
// ------------------------------------------------------------------------------
// TCH: tracers per Thomas chunk; CPB: columns per block (one column per LB
// lanes). The Thomas sweeps of all CPB x TCH (column, tracer) pairs of a block
// run in one wave, one lane per pair, while the other waves wait: a serial
// sweep then costs one wave's issue slots for CPB columns instead of one per
// column. Both are launch-shape choices; the arithmetic is the same for all.
// XF bit 0 (use_wsplit): the non-FCT tracers' implicit diffusion also takes
// the implicit part w_i of the split W (second set of rows). XF bit 1: the
// explicit diffusion increments (Ki, Redi, i_vert_diff off) and the Redi K33
// terms are compiled in. Both are launched only when in use.
template <int LB, int TCH, int CPB, int XF>
__global__ void __launch_bounds__(LB * CPB) __attribute__((amdgpu_waves_per_eu(6))) kD_node(DevMesh m, DevDyn d, DevTracers t, double dt, int clamp_tracer,
                                                    int tpb)
{
    const int w = threadIdx.x / LB;
    const int j0 = xcd_column(m.N2) * CPB + w;
    const bool valid = j0 < m.N2;
    const int n = m.node_ord[valid ? j0 : m.N2 - 1];
    // LDS reuse: Z_n lives in sAv's rows until the tracer loop starts, and the
    // Thomas sweep's c' rows of the first two tracers of a chunk live in sRp
    // and sRm, which are dead between a chunk's flux phase and its sweep
    __shared__ double sRpA[CPB][LB + 2], sRmA[CPB][LB + 2], sAvA[CPB][LB + 2];
    __shared__ double saA[CPB][LB + 2], sbA[CPB][LB + 2], scA[CPB][LB + 2];
    __shared__ double sTsA[CPB][TCH][LB + 2], stpA[CPB][TCH][LB + 2], scpX[CPB][TCH > 2 ? TCH - 2 : 1][LB + 2];
    __shared__ int sNz[CPB][2];
    constexpr bool WS = (XF & 1) != 0, DX = (XF & 2) != 0;
    constexpr int NW = WS ? LB + 2 : 1;
    __shared__ double swaA[CPB][NW], swbA[CPB][NW], swcA[CPB][NW];
    double *sRp = sRpA[w], *sRm = sRmA[w], *sAv = sAvA[w];
    double *sa = saA[w], *sb = sbA[w], *sc = scA[w], *sZ = sAvA[w];
    const int tid = threadIdx.x % LB, k = tid + 1;
    const int L = m.L, NL = m.NL;
    const int nzmin = valid ? m.ulevn[n] : 1, nzmax = valid ? m.nlevn[n] : 0;
    if (tid == 0) { sNz[w][0] = nzmin; sNz[w][1] = nzmax; }
    const bool act = (k >= nzmin && k <= nzmax - 1);
    const size_t ci = (size_t)n * L + (k <= L ? k - 1 : L - 1);
    const NodeEdges E = load_node_edges(m, n);
    double hn = 1.0, hh = 0.0, av_ = 1.0, arv = 0.0;
    if (act) {
        hn = m.hnode_new[ci];
        hh = m.hnode[ci];
        av_ = m.areasvol[(size_t)n * NL + k - 1];
        arv = m.area[(size_t)n * NL + k - 1];
    }
    // Z_n of the implicit diffusion, rebuilt from hnode_new when the geometry is set
    if (act) sZ[tid] = m.zn_vd[ci];
    __syncthreads();
    // tridiagonal coefficients (oce_ale_tracer.F90:1098-1175): tracer independent
    if (act) {
        const double *kv = d.Kv + (size_t)n * NL;
        const double zinv = 1.0 * dt;
        double a = 0.0, b, c;
        // Redi: K33 = Kv + Ki S3^2 at the interfaces (oce_ale_tracer.F90:1105-1108, 1138-1144, 1181-1184)
        double tya = 0.0, tyc = 0.0;
        const bool redi = DX && d.redi != 0;
        if (redi) {
            const double *zbn = m.zbn_vd + (size_t)n * NL;
            const double *S = d.slope + (size_t)n * L * 3, *ki = d.Ki + (size_t)n * L;
            const double s3 = S[(k - 1) * 3 + 2];
            if (k <= nzmax - 2) {   // interface below layer k
                const double zinv2 = 1.0 / (sZ[k - 1] - sZ[k]);
                const double s3b = S[k * 3 + 2];
                tyc = (zbn[k] - sZ[k]) * zinv2 * (s3 * s3) * ki[k - 1] +
                      (sZ[k - 1] - zbn[k]) * zinv2 * (s3b * s3b) * ki[k];
            }
            if (k > nzmin) {   // interface above layer k
                const double zinv1 = 1.0 / (sZ[k - 2] - sZ[k - 1]);
                const double s3a = S[(k - 2) * 3 + 2];
                tya = (zbn[k - 1] - sZ[k - 1]) * zinv1 * (s3a * s3a) * ki[k - 2] +
                      (sZ[k - 2] - zbn[k - 1]) * zinv1 * (s3 * s3) * ki[k - 1];
            }
        }
        if (k == nzmin) {
            const double zinv2 = 1.0 / (sZ[k - 1] - sZ[k]);
            c = redi ? -(kv[k] + tyc) * zinv2 * zinv * m.area[(size_t)n * NL + k] / av_
                     : -(kv[k]) * zinv2 * zinv * m.area[(size_t)n * NL + k] / av_;
            b = -c + hn;
        } else if (k <= nzmax - 2) {
            const double zinv1 = 1.0 / (sZ[k - 2] - sZ[k - 1]);
            const double zinv2 = 1.0 / (sZ[k - 1] - sZ[k]);
            if (redi) {
                a = -(kv[k - 1] + tya) * zinv1 * zinv * (arv / av_);
                c = -(kv[k] + tyc) * zinv2 * zinv * m.area[(size_t)n * NL + k] / av_;
            } else {
                a = -(kv[k - 1]) * zinv1 * zinv * (arv / av_);
                c = -(kv[k]) * zinv2 * zinv * m.area[(size_t)n * NL + k] / av_;
            }
            b = -a - c + hn;
        } else {
            const double zinv1 = 1.0 / (sZ[k - 2] - sZ[k - 1]);
            a = redi ? -(kv[k - 1] + tya) * zinv1 * zinv * (arv / av_) : -(kv[k - 1]) * zinv1 * zinv * (arv / av_);
            c = 0.0;
            b = -a + hn;
        }
        sa[tid] = a;
        sb[tid] = b;
        sc[tid] = c;
        if (WS) {   // oce_ale_tracer.F90:1118-1129, 1158-1172, 1194-1202
            const double *wi = d.w_i + (size_t)n * NL;
            double v_adv = zinv * (arv / av_);
            if (k == nzmin) {
                b = b + wi[k - 1] * v_adv;
            } else {
                a = a + dmn(0.0, wi[k - 1]) * v_adv;
                b = b + dmx(0.0, wi[k - 1]) * v_adv;
            }
            if (k <= nzmax - 2) {
                v_adv = zinv * m.area[(size_t)n * NL + k] / av_;
                b = b - dmn(0.0, wi[k]) * v_adv;
                c = c - dmx(0.0, wi[k]) * v_adv;
            }
            swaA[w][tid] = a;
            swbA[w][tid] = b;
            swcA[w][tid] = c;
        }
    }
    const int trb = blockIdx.y * tpb, tre = min(t.ntr, trb + tpb);
    for (int t0 = trb; t0 < tre; t0 += TCH) {
        const int nch = min(TCH, tre - t0);
        for (int c = 0; c < nch; ++c) {
            const int tr = t0 + c;
            double *T = t.T + t.sN * tr;
            double *Told = t.Told + t.sOld * tr;
            const int abo = t.abo[tr];
            const bool fct = t.lim[tr] == LIM_FCT;
            const double *av = t.a_v + t.sNI * tr + (size_t)n * NL;
            __syncthreads();
            if (fct && k <= L) {
                sRp[tid] = t.rp[t.sN * tr + ci];
                sRm[tid] = t.rm[t.sN * tr + ci];
            }
            __syncthreads();
            // limited vertical antidiffusive flux at interface k (oce_adv_tra_fct.F90:425-456)
            if (k >= nzmin && k <= nzmax) {
                const double flux = av[k - 1];
                double v = flux;
                if (fct && k <= nzmax - 1) {
                    double ae = 1.0;
                    if (k == nzmin) {
                        if (flux >= 0.0) ae = dmn(ae, sRp[tid]);
                        else ae = dmn(ae, sRm[tid]);
                    } else {
                        if (flux >= 0.0) {
                            ae = dmn(ae, sRm[tid - 1]);
                            ae = dmn(ae, sRp[tid]);
                        } else {
                            ae = dmn(ae, sRp[tid - 1]);
                            ae = dmn(ae, sRm[tid]);
                        }
                    }
                    v = ae * flux;
                }
                sAv[tid] = v;
            }
            __syncthreads();
            double Tk = 0.0, v = 0.0, h = 0.0;
            if (act) {
                Tk = T[ci];
                if (fct) v = v - Tk * hh + t.fct_lo[t.sN * tr + ci] * hn;
                v = v + (sAv[tid] - sAv[tid + 1]) * dt / av_;
            }
            const double *ah = t.a_h + t.sE * tr;
            const double rpn = (act && fct) ? sRp[tid] : 1.0, rmn = (act && fct) ? sRm[tid] : 1.0;
            // 128 lanes (columns of 65-127 layers): edges in batches of two,
            // the batch's loads (clamped to the column) issued together, then
            // added in ascending edge order (at 1.3 M columns of up to 70
            // layers and 10 tracers: 74 -> 66 ms per launch). At 64 lanes the
            // extra registers cost more than the overlap gains (0.76 -> 0.79 ms
            // at 127 k columns), so there the loop takes one edge at a time.
            if constexpr (LB == 128) {
                constexpr int KD_NG = 2;
                const int kk = min(k, L) - 1;
                for (int j0 = 0; j0 < E.deg; j0 += KD_NG) {
                    double fl_[KD_NG], rpo_[KD_NG], rmo_[KD_NG];
                    int q_[KD_NG];
                    bool ok_[KD_NG];
#pragma unroll
                    for (int u = 0; u < KD_NG; ++u) {
                        const bool real = j0 + u < E.deg;
                        const int j = real ? j0 + u : 0;
                        const int q = rl(E.q, j), mo = rl(E.other, j), a0 = rl(E.nu12, j), a1 = rl(E.nl12, j);
                        ok_[u] = real && act && k >= a0 && k <= a1;
                        q_[u] = q;
                        fl_[u] = ah[(size_t)(q >> 1) * L + kk];
                        if (fct) {
                            const size_t oi = (size_t)mo * L + kk;
                            rpo_[u] = t.rp[t.sN * tr + oi];
                            rmo_[u] = t.rm[t.sN * tr + oi];
                        }
                    }
#pragma unroll
                    for (int u = 0; u < KD_NG; ++u) {
                        if (!ok_[u]) continue;
                        const int q = q_[u];
                        double fl = fl_[u];
                        if (fct) fl = limit_edge_flux(fl, q, rpn, rmn, rpo_[u], rmo_[u]);
                        if (q & 1) h = h - fl * dt / av_;
                        else h = h + fl * dt / av_;
                    }
                }
            } else {
                for (int j = 0; j < E.deg; ++j) {
                    const int q = rl(E.q, j), mo = rl(E.other, j), a0 = rl(E.nu12, j), a1 = rl(E.nl12, j);
                    if (!(act && k >= a0 && k <= a1)) continue;
                    double fl = ah[(size_t)(q >> 1) * L + k - 1];
                    if (fct) {
                        const size_t oi = (size_t)mo * L + k - 1;
                        fl = limit_edge_flux(fl, q, rpn, rmn, t.rp[t.sN * tr + oi], t.rm[t.sN * tr + oi]);
                    }
                    if (q & 1) h = h - fl * dt / av_;
                    else h = h + fl * dt / av_;
                }
            }
            double del = 0.0 + h + v;
            // explicit diffusion in diff_tracers_ale's order (oce_ale_tracer.F90:595-626):
            // edge fluxes of diff_part_hor_redi ascending, then the vertical parts
            if (DX && d.hdiff) {
                const double *hd = t.hd + t.sE * tr;
                for (int j = 0; j < E.deg; ++j) {
                    const int q = rl(E.q, j), a0 = rl(E.nu12, j), a1 = rl(E.nl12, j);
                    if (!(act && k >= a0 && k <= a1)) continue;
                    const double c = hd[(size_t)(q >> 1) * L + k - 1];
                    const double r = (q & 1) ? 0.0 - c : 0.0 + c;
                    del = del + r * dt / av_;
                }
            }
            if (act) {
                if (DX && !t.ivd[tr]) del = del + t.ve[t.sN * tr + ci];
                if (DX && d.redi) del = del + t.vr[t.sN * tr + ci];
                del = del + Tk * (hh - hn);
                sTsA[w][c][tid] = Tk + del / hn;
                // valuesold rotation (oce_tracer_mod.F90:94-123)
                if (abo == 2) {
                    Told[ci] = Tk;
                } else {
                    Told[2 * ci + 1] = Told[2 * ci];
                    Told[2 * ci] = Tk;
                }
            }
        }
        __syncthreads();
        // right-hand sides (oce_ale_tracer.F90:1206-1229, 1377-1383), all tracers of the chunk
        for (int c = 0; c < nch; ++c) {
            const int tr = t0 + c;
            if (!t.ivd[tr] || !act) continue;
            const double dz = hn;
            const double *ts = sTsA[w][c];
            const bool wim = WS && t.lim[tr] != LIM_FCT;
            const double *ra = wim ? swaA[w] : sa, *rb = wim ? swbA[w] : sb, *rc = wim ? swcA[w] : sc;
            double r;
            if (k == nzmin) r = -(rb[tid] - dz) * ts[tid] - rc[tid] * ts[tid + 1];
            else if (k <= nzmax - 2) r = -ra[tid] * ts[tid - 1] - (rb[tid] - dz) * ts[tid] - rc[tid] * ts[tid + 1];
            else r = -ra[tid] * ts[tid - 1] - (rb[tid] - dz) * ts[tid];
            if (bc_every_layer(t.id[tr], d)) r = r + bc_surface_dev(t.id[tr], dt, d, m.N2, L, n, k, ts[tid]);
            else if (k == nzmin) r = r + bc_surface_dev(t.id[tr], dt, d, m.N2, L, n, nzmin, ts[tid]);
            stpA[w][c][tid] = r;
        }
        __syncthreads();
        // Thomas sweeps (oce_ale_tracer.F90:1406-1430): one lane of the first wave
        // per (column, tracer) pair; the pivot chain (cp, m) is evaluated in the
        // same loop as tp so that the two recurrences overlap
        if (threadIdx.x < CPB * nch) {
            const int wc = threadIdx.x / nch, c = threadIdx.x % nch;
            const int z0 = sNz[wc][0], z1 = sNz[wc][1];
            if (z1 - 1 >= z0 && t.ivd[t0 + c]) {
                const bool wim = WS && t.lim[t0 + c] != LIM_FCT;
                const double *a_ = wim ? swaA[wc] : saA[wc], *b_ = wim ? swbA[wc] : sbA[wc], *c_ = wim ? swcA[wc] : scA[wc];
                double *tp_ = stpA[wc][c], *ts = sTsA[wc][c];
                double *cp_ = c == 0 ? sRpA[wc] : (c == 1 ? sRmA[wc] : scpX[wc][c >= 2 ? c - 2 : 0]);
                double cp = c_[z0 - 1] / b_[z0 - 1];
                double tp = tp_[z0 - 1] / b_[z0 - 1];
                cp_[z0 - 1] = cp;
                tp_[z0 - 1] = tp;
                for (int nz = z0 + 1; nz <= z1 - 1; ++nz) {
                    const double an = a_[nz - 1];
                    const double mm = b_[nz - 1] - cp * an;
                    cp = c_[nz - 1] / mm;
                    tp = (tp_[nz - 1] - tp * an) / mm;
                    cp_[nz - 1] = cp;
                    tp_[nz - 1] = tp;
                }
                double x = tp_[z1 - 2];
                ts[z1 - 2] = ts[z1 - 2] + x;
                for (int nz = z1 - 2; nz >= z0; --nz) {
                    x = tp_[nz - 1] - cp_[nz - 1] * x;
                    ts[nz - 1] = ts[nz - 1] + x;
                }
            }
        }
        __syncthreads();
        if (act) {
            for (int c = 0; c < nch; ++c) {
                const int tr = t0 + c;
                double v = sTsA[w][c][tid];
                if (tr == clamp_tracer) {
                    if (v > 45.0) v = 45.0;
                    if (v < 3.0) v = 3.0;
                }
                t.T[t.sN * tr + ci] = v;
            }
        }
    }
}

// ------------------------------------------------------------------------------
// K_D with one wave per column (plain step: no W split, no explicit diffusion)
// ------------------------------------------------------------------------------
// One wave per node column, S levels per lane (level k = lane + 1 + 64 s), so
// every exchange between levels of a column stays inside one wave and needs
// no workgroup barrier. The tracers of the block are taken in chunks of TCH.
// SW: CPB column waves plus one sweep wave per block. In phase p the column
//     waves compute the limited fluxes, T* and the tridiagonal right-hand
//     sides of chunk p into one of two LDS buffers, while the sweep wave runs
//     the Thomas sweeps of chunk p-1 (one lane per (column, tracer) pair) from
//     the other buffer and stores that chunk's T^{n+1}; the serial sweeps
//     overlap the level-parallel work of the next chunk.
// !SW: CPB independent column waves, each of which sweeps and stores its own
//     column after each chunk.
// Same arithmetic as kD_node, so the results are bit-identical to it. Which of
// the three is fastest depends on the workload (the autotuner picks): at
// 1.3 M columns of up to 70 levels and 10 tracers the sweep wave gives
// 63.5 ms per launch against 66.8 ms lock-step; at 127 k columns of up to
// 47 levels and 2 tracers (one chunk, nothing to overlap) the lock-step
// kernel is fastest (0.76 ms, against 0.84 ms wave-local).


// A column's values one level up and one level down the wave: lane j, slot s
// holds level j + 1 + 64 s (the lane below, or across the slot boundary)
template <int S>
__device__ __forceinline__ void level_up(const double (&x)[S], double (&y)[S], int lane)
{
    y[0] = __shfl_up(x[0], 1);   // lane 0 gets its own value (level 0 does not exist)
    if constexpr (S == 2) {
        const double last0 = __shfl(x[0], 63);
        const double u1 = __shfl_up(x[1], 1);
        y[1] = lane == 0 ? last0 : u1;
    }
}
template <int S>
__device__ __forceinline__ void level_down(const double (&x)[S], double (&y)[S], int lane)
{
    const double d0 = __shfl_down(x[0], 1);
    if constexpr (S == 2) {
        const double first1 = __shfl(x[1], 0);
        y[0] = lane == 63 ? first1 : d0;
        y[1] = __shfl_down(x[1], 1);
    } else {
        y[0] = d0;
    }
}

// A one-wave block of short rows is held to 96 VGPRs (5 waves per SIMD; it
// spills 40 bytes per lane): 54.1 -> 52.8 ms at 1.3 M columns of 70 levels
template <int S, int TCH, int CPB, bool SW, int RS = 64 * S + 2, bool SH = false>
__global__ void __launch_bounds__(64 * (CPB + (SW ? 1 : 0)))
__attribute__((amdgpu_waves_per_eu((!SW && CPB == 1 && S == 2 && RS <= KDP_SHORT_NL + 2) ? 5 : 1, 8)))
kDp_node(DevMesh m, DevDyn d, DevTracers t, double dt, int clamp_tracer, int tpb)
{
    // RS: LDS row of at least NL + 2 doubles (index k - 1 of levels and
    // interfaces 1..NL, plus a guard); every LDS access is within k <= NL
    constexpr int NB = SW ? 2 : 1;   // chunk buffers
    // SH: R+- and the limited vertical fluxes, read one level away only, move
    // between lanes by shuffles instead of LDS rows
    __shared__ double sRpA[SH ? 1 : CPB][SH ? 1 : RS], sRmA[SH ? 1 : CPB][SH ? 1 : RS], sAvA[SH ? 1 : CPB][SH ? 1 : RS];
    __shared__ double saA[CPB][RS], sbA[CPB][RS], scA[CPB][RS], scpA[CPB][RS];
    __shared__ double sTsA[NB][CPB][TCH][RS], stpA[NB][CPB][TCH][RS];
    __shared__ int sNz[CPB][2];
    const int w = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const bool sweeper = SW && w == CPB;
    const int L = m.L, NL = m.NL;
    const int col0 = xcd_column(m.N2) * CPB;
    const int trb = blockIdx.y * tpb, tre = min(t.ntr, trb + tpb);
    const int nchunk = (tre - trb + TCH - 1) / TCH;

    // ---- column waves: geometry and the tracer-independent tridiagonal rows
    const int wc_ = sweeper ? 0 : w;
    const int j0 = col0 + wc_;
    const bool valid = !sweeper && j0 < m.N2;
    const int n = m.node_ord[j0 < m.N2 ? j0 : m.N2 - 1];
    const int nzmin = valid ? m.ulevn[n] : 1, nzmax = valid ? m.nlevn[n] : 0;
    double *sRp = sRpA[SH ? 0 : wc_], *sRm = sRmA[SH ? 0 : wc_], *sAv = sAvA[SH ? 0 : wc_];
    double *sZ = SH ? sTsA[0][wc_][0] : sAvA[wc_];   // a row that is dead before the first chunk
    double *sa = saA[wc_], *sb = sbA[wc_], *sc = scA[wc_];
    int kk_[S];
    bool act_[S];
    size_t ci_[S];
    double hn_[S], hh_[S], av_[S], arv_[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int k = lane + 1 + 64 * s;
        kk_[s] = k;
        act_[s] = k >= nzmin && k <= nzmax - 1;
        ci_[s] = (size_t)n * L + (k <= L ? k - 1 : L - 1);
        hn_[s] = 1.0; hh_[s] = 0.0; av_[s] = 1.0; arv_[s] = 0.0;
        if (act_[s]) {
            hn_[s] = m.hnode_new[ci_[s]];
            hh_[s] = m.hnode[ci_[s]];
            av_[s] = m.areasvol[(size_t)n * NL + k - 1];
            arv_[s] = m.area[(size_t)n * NL + k - 1];
        }
    }
    NodeEdges E;
    if (!sweeper) {
        E = load_node_edges(m, n);
        if (lane == 0) { sNz[w][0] = nzmin; sNz[w][1] = nzmax; }
#pragma unroll
        for (int s = 0; s < S; ++s)
            if (act_[s]) sZ[kk_[s] - 1] = m.zn_vd[ci_[s]];
        wave_sync();
        // tridiagonal coefficients (oce_ale_tracer.F90:1098-1175)
#pragma unroll
        for (int s = 0; s < S; ++s) {
            if (!act_[s]) continue;
            const int k = kk_[s];
            const double *kv = d.Kv + (size_t)n * NL;
            const double zinv = 1.0 * dt;
            double a = 0.0, b, c;
            if (k == nzmin) {
                const double zinv2 = 1.0 / (sZ[k - 1] - sZ[k]);
                c = -(kv[k]) * zinv2 * zinv * m.area[(size_t)n * NL + k] / av_[s];
                b = -c + hn_[s];
            } else if (k <= nzmax - 2) {
                const double zinv1 = 1.0 / (sZ[k - 2] - sZ[k - 1]);
                const double zinv2 = 1.0 / (sZ[k - 1] - sZ[k]);
                a = -(kv[k - 1]) * zinv1 * zinv * (arv_[s] / av_[s]);
                c = -(kv[k]) * zinv2 * zinv * m.area[(size_t)n * NL + k] / av_[s];
                b = -a - c + hn_[s];
            } else {
                const double zinv1 = 1.0 / (sZ[k - 2] - sZ[k - 1]);
                a = -(kv[k - 1]) * zinv1 * zinv * (arv_[s] / av_[s]);
                c = 0.0;
                b = -a + hn_[s];
            }
            sa[k - 1] = a;
            sb[k - 1] = b;
            sc[k - 1] = c;
        }
    }
    if (SW) __syncthreads();

    // The forward-elimination pivots of the columns (oce_ale_tracer.F90:1406-1414):
    // m = b - c'a and c' = c / m depend on the column only. With the sweep
    // wave, its first CPB lanes compute the c' rows while the column waves
    // work on the first chunk; every sweep then forms m again from b, a and
    // the c' above (the same two operations) and divides once per level.
    auto pivots = [&]() {
        if (lane < CPB) {
            const int wc = lane, z0 = sNz[wc][0], z1 = sNz[wc][1];
            if (z1 - 1 >= z0) {
                const double *a_ = saA[wc], *b_ = sbA[wc], *c_ = scA[wc];
                double *cp_ = scpA[wc];
                double cp = c_[z0 - 1] / b_[z0 - 1];
                cp_[z0 - 1] = cp;
                for (int nz = z0 + 1; nz <= z1 - 1; ++nz) {
                    const double mm = b_[nz - 1] - cp * a_[nz - 1];
                    cp = c_[nz - 1] / mm;
                    cp_[nz - 1] = cp;
                }
            }
        }
    };

    // Without the sweep wave, the first sweep of a column writes its c' row
    // (every lane of the column the same values), and the later chunks' sweeps
    // reuse it as the sweep wave does
    bool piv_ok = false;

    // Thomas sweeps of chunk q (oce_ale_tracer.F90:1406-1430), one lane per
    // (column, tracer) of the columns wlo..wlo+nw-1, then T^{n+1} of those
    // columns with the salinity clamp
    auto sweep_store = [&](int q, int wlo, int nw) {
        const int t0 = trb + q * TCH, nch = min(TCH, tre - t0), bf = SW ? (q & 1) : 0;
        if (lane < nw * nch) {
            const int wc = wlo + lane / nch, c = lane % nch;
            const int z0 = sNz[wc][0], z1 = sNz[wc][1];
            if (z1 - 1 >= z0 && t.ivd[t0 + c]) {
                const double *a_ = saA[wc], *b_ = sbA[wc], *c_ = scA[wc];
                double *tp_ = stpA[bf][wc][c], *ts = sTsA[bf][wc][c], *cp_ = scpA[wc];
                double tp = tp_[z0 - 1] / b_[z0 - 1];
                tp_[z0 - 1] = tp;
                if (SW || piv_ok) {
                    double cpa = cp_[z0 - 1];   // c' of the level above
                    for (int nz = z0 + 1; nz <= z1 - 1; ++nz) {
                        const double an = a_[nz - 1];
                        const double mm = b_[nz - 1] - cpa * an;
                        tp = (tp_[nz - 1] - tp * an) / mm;
                        cpa = cp_[nz - 1];
                        tp_[nz - 1] = tp;
                    }
                } else {
                    double cp = c_[z0 - 1] / b_[z0 - 1];
                    cp_[z0 - 1] = cp;
                    for (int nz = z0 + 1; nz <= z1 - 1; ++nz) {
                        const double an = a_[nz - 1];
                        const double mm = b_[nz - 1] - cp * an;
                        cp = c_[nz - 1] / mm;
                        tp = (tp_[nz - 1] - tp * an) / mm;
                        cp_[nz - 1] = cp;
                        tp_[nz - 1] = tp;
                    }
                }
                double x = tp_[z1 - 2];
                ts[z1 - 2] = ts[z1 - 2] + x;
                for (int nz = z1 - 2; nz >= z0; --nz) {
                    x = tp_[nz - 1] - cp_[nz - 1] * x;
                    ts[nz - 1] = ts[nz - 1] + x;
                }
            }
        }
        wave_sync();
        for (int wc = wlo; wc < wlo + nw; ++wc) {
            if (col0 + wc >= m.N2) break;
            const int nn = m.node_ord[col0 + wc];
            const int z0 = sNz[wc][0], z1 = sNz[wc][1];
            for (int c = 0; c < nch; ++c) {
                const int tr = t0 + c;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int k = lane + 1 + 64 * s;
                    if (k < z0 || k > z1 - 1) continue;
                    double v = sTsA[bf][wc][c][k - 1];
                    if (tr == clamp_tracer) {
                        if (v > 45.0) v = 45.0;
                        if (v < 3.0) v = 3.0;
                    }
                    t.T[t.sN * tr + (size_t)nn * L + k - 1] = v;
                }
            }
        }
    };

    for (int p = 0; p <= nchunk; ++p) {
        if (!sweeper && p < nchunk) {
            // ---- chunk p: limited fluxes, flux2dtracer, T*, right-hand sides
            const int t0 = trb + p * TCH, nch = min(TCH, tre - t0), bf = SW ? (p & 1) : 0;
            for (int c = 0; c < nch; ++c) {
                const int tr = t0 + c;
                double *T = t.T + t.sN * tr;
                double *Told = t.Told + t.sOld * tr;
                const int abo = t.abo[tr];
                const bool fct = t.lim[tr] == LIM_FCT;
                const double *av = t.a_v + t.sNI * tr + (size_t)n * NL;
                double rp_[S], rm_[S], rpu_[S], rmu_[S], avk_[S], avd_[S];
                if constexpr (SH) {
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        rp_[s] = 1.0; rm_[s] = 1.0;
                        if (fct && act_[s]) {
                            rp_[s] = t.rp[t.sN * tr + ci_[s]];
                            rm_[s] = t.rm[t.sN * tr + ci_[s]];
                        }
                    }
                    level_up<S>(rp_, rpu_, lane);
                    level_up<S>(rm_, rmu_, lane);
                } else {
                    wave_sync();
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        const int k = kk_[s];
                        if (fct && act_[s]) {   // the limiters read R+- of the wet levels only
                            sRp[k - 1] = t.rp[t.sN * tr + ci_[s]];
                            sRm[k - 1] = t.rm[t.sN * tr + ci_[s]];
                        }
                    }
                    wave_sync();
                }
                // limited vertical antidiffusive flux at interface k (oce_adv_tra_fct.F90:425-456)
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int k = kk_[s];
                    avk_[s] = 0.0;
                    if (k >= nzmin && k <= nzmax) {
                        const double flux = av[k - 1];
                        double v = flux;
                        if (fct && k <= nzmax - 1) {
                            const double rpk = SH ? rp_[s] : sRp[k - 1], rmk = SH ? rm_[s] : sRm[k - 1];
                            double ae = 1.0;
                            if (k == nzmin) {
                                if (flux >= 0.0) ae = dmn(ae, rpk);
                                else ae = dmn(ae, rmk);
                            } else {
                                if (flux >= 0.0) {
                                    ae = dmn(ae, SH ? rmu_[s] : sRm[k - 2]);
                                    ae = dmn(ae, rpk);
                                } else {
                                    ae = dmn(ae, SH ? rpu_[s] : sRp[k - 2]);
                                    ae = dmn(ae, rmk);
                                }
                            }
                            v = ae * flux;
                        }
                        if (SH) avk_[s] = v;
                        else sAv[k - 1] = v;
                    }
                }
                if constexpr (SH) level_down<S>(avk_, avd_, lane);
                else wave_sync();
                double Tk_[S], v_[S], h_[S], rpn_[S], rmn_[S];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int k = kk_[s];
                    Tk_[s] = 0.0; v_[s] = 0.0; h_[s] = 0.0;
                    if (act_[s]) {
                        Tk_[s] = T[ci_[s]];
                        if (fct) v_[s] = v_[s] - Tk_[s] * hh_[s] + t.fct_lo[t.sN * tr + ci_[s]] * hn_[s];
                        v_[s] = v_[s] + (SH ? avk_[s] - avd_[s] : sAv[k - 1] - sAv[k]) * dt / av_[s];
                    }
                    rpn_[s] = (act_[s] && fct) ? (SH ? rp_[s] : sRp[k - 1]) : 1.0;
                    rmn_[s] = (act_[s] && fct) ? (SH ? rm_[s] : sRm[k - 1]) : 1.0;
                }
                const double *ah = t.a_h + t.sE * tr;
                for (int j = 0; j < E.deg; ++j) {
                    const int q = rl(E.q, j), mo = rl(E.other, j), a0 = rl(E.nu12, j), a1 = rl(E.nl12, j);
                    double fl_[S], rpo_[S], rmo_[S];
                    bool ok_[S];
#pragma unroll
                    for (int s = 0; s < S; ++s) {   // the slots' loads together
                        const int k = kk_[s];
                        ok_[s] = act_[s] && k >= a0 && k <= a1;
                        fl_[s] = 0.0; rpo_[s] = 1.0; rmo_[s] = 1.0;
                        if (ok_[s]) {
                            fl_[s] = ah[(size_t)(q >> 1) * L + k - 1];
                            if (fct) {
                                const size_t oi = (size_t)mo * L + k - 1;
                                rpo_[s] = t.rp[t.sN * tr + oi];
                                rmo_[s] = t.rm[t.sN * tr + oi];
                            }
                        }
                    }
#pragma unroll
                    for (int s = 0; s < S; ++s) {
                        if (!ok_[s]) continue;
                        double fl = fl_[s];
                        if (fct) fl = limit_edge_flux(fl, q, rpn_[s], rmn_[s], rpo_[s], rmo_[s]);
                        if (q & 1) h_[s] = h_[s] - fl * dt / av_[s];
                        else h_[s] = h_[s] + fl * dt / av_[s];
                    }
                }
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (!act_[s]) continue;
                    const size_t ci = ci_[s];
                    double del = 0.0 + h_[s] + v_[s];
                    del = del + Tk_[s] * (hh_[s] - hn_[s]);
                    sTsA[bf][w][c][kk_[s] - 1] = Tk_[s] + del / hn_[s];
                    // valuesold rotation (oce_tracer_mod.F90:94-123)
                    if (abo == 2) {
                        Told[ci] = Tk_[s];
                    } else {
                        Told[2 * ci + 1] = Told[2 * ci];
                        Told[2 * ci] = Tk_[s];
                    }
                }
            }
            wave_sync();
            // right-hand sides (oce_ale_tracer.F90:1206-1229, 1377-1383)
            for (int c = 0; c < nch; ++c) {
                const int tr = t0 + c;
                if (!t.ivd[tr]) continue;
                const double *ts = sTsA[bf][w][c];
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    if (!act_[s]) continue;
                    const int k = kk_[s], i = k - 1;
                    const double dz = hn_[s];
                    double r;
                    if (k == nzmin) r = -(sb[i] - dz) * ts[i] - sc[i] * ts[i + 1];
                    else if (k <= nzmax - 2) r = -sa[i] * ts[i - 1] - (sb[i] - dz) * ts[i] - sc[i] * ts[i + 1];
                    else r = -sa[i] * ts[i - 1] - (sb[i] - dz) * ts[i];
                    if (bc_every_layer(t.id[tr], d)) r = r + bc_surface_dev(t.id[tr], dt, d, m.N2, L, n, k, ts[i]);
                    else if (k == nzmin) r = r + bc_surface_dev(t.id[tr], dt, d, m.N2, L, n, nzmin, ts[i]);
                    stpA[bf][w][c][i] = r;
                }
            }
            if (!SW) {   // this wave sweeps and stores its own column
                wave_sync();
                sweep_store(p, w, 1);
                for (int c = 0; c < nch && !piv_ok; ++c) piv_ok = t.ivd[t0 + c] != 0;
            }
        } else if (sweeper && p >= 1) {
            sweep_store(p - 1, 0, CPB);
        } else if (sweeper) {
            pivots();
        }
        if (SW) __syncthreads();
    }
}

// Z levels of diff_ver_part_impl_ale rebuilt bottom-up from hnode_new and
// zbar_n_bot (oce_ale_tracer.F90:1088-1096), one thread per column
__global__ void k_zn_vdiff(DevMesh m, double *zn, double *zbn)
{
    const int n = blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= m.N2) return;
    const int L = m.L, nzmin = m.ulevn[n], nzmax = m.nlevn[n];
    const double *hn = m.hnode_new + (size_t)n * L;
    double *z = zn + (size_t)n * L, *zi = zbn + (size_t)n * m.NL;
    double zb = m.zbar_n_bot[n];
    zi[nzmax - 1] = zb;
    z[nzmax - 2] = zb + hn[nzmax - 2] / 2.0;
    for (int nz = nzmax - 1; nz >= nzmin + 1; --nz) {
        zb = zb + hn[nz - 1];
        zi[nz - 1] = zb;
        z[nz - 2] = zb + hn[nz - 2] / 2.0;
    }
}

// ==============================================================================
// host side
// ==============================================================================
int fgpu_compute_zn(fgpu_context *c)
{
    if (!c->m.zn_vd) {
        double *p;
        int rc = fgpu_alloc(c, (void **)&p, sizeof(double) * (size_t)c->N2 * (c->nl - 1));
        if (rc) return rc;
        FGPU_HIP(hipMemset(p, 0, sizeof(double) * (size_t)c->N2 * (c->nl - 1)));
        c->m.zn_vd = p;
        if ((rc = fgpu_alloc(c, (void **)&p, sizeof(double) * (size_t)c->N2 * c->nl))) return rc;
        FGPU_HIP(hipMemset(p, 0, sizeof(double) * (size_t)c->N2 * c->nl));
        c->m.zbn_vd = p;
    }
    hipLaunchKernelGGL(k_zn_vdiff, dim3((c->N2 + 255) / 256), dim3(256), 0, c->stream, c->m, (double *)c->m.zn_vd,
                       (double *)c->m.zbn_vd);
    FGPU_HIP(hipGetLastError());
    FGPU_HIP(hipStreamSynchronize(c->stream));
    return FGPU_OK;
}
// Extra terms compiled into K_D: bit 0 the W split (w_i rows of the implicit
// diffusion for non-FCT tracers), bit 1 explicit diffusion increments (Ki,
// Redi, i_vert_diff off) and the Redi K33 rows. Only the plain step (0) has
// the one-wave-per-column variants.
int fgpu_kd_extras(const fgpu_context *c)
{
    bool dx = c->d.hdiff || c->d.redi;
    for (int v : c->h_ivd) dx = dx || !v;
    return (c->d.wsplit ? 1 : 0) | (dx ? 2 : 0);
}

// The Thomas chunk and columns per block K_D runs with, for the requested
// ones: several columns per block take chunks of 2 (or 4 at four 64-lane
// columns, or two columns of up to 128 levels); several columns of more than
// 64 levels need the sweep-wave K_D.
void fgpu_kd_shape(const fgpu_context *c, int *tch_out, int *cpb_out)
{
    int tch = (c->step_tch == 4 || c->step_tch == 5 || c->step_tch == 8) ? c->step_tch : 2;
    int cpb = c->step_cpb;
    if (fgpu_kd_extras(c)) {   // the W split or explicit diffusion terms: one lock-step shape
        tch = 2;
        cpb = 1;
    } else if (c->m.NL <= 64) {
        if (cpb == 8) tch = 2;
        else if (cpb == 4) tch = tch == 4 ? 4 : 2;
        else cpb = 1;
    } else if (c->step_kdp && cpb >= 3 && c->m.NL <= KDP_SHORT_NL) {
        cpb = cpb >= 4 ? 4 : 3;
        tch = 2;
    } else if (c->step_kdp && cpb >= 2) {
        cpb = 2;
        tch = tch == 4 ? 4 : 2;
    } else {
        cpb = 1;
    }
    *tch_out = tch;
    *cpb_out = cpb;
}

// K_D with one wave per column (S levels per lane: 1 up to 64 interfaces, 2 up
// to 128), with a sweep wave per block (SW) or each column wave sweeping its
// own column, for the Thomas chunk and columns per block of the launch shape
template <int S, int TCH, int CPB, bool SW, int RS = 64 * S + 2, bool SH = false>
static void launch_kdp(fgpu_context *c, unsigned ny, double dt, int clamp, int tpb)
{
    const DevMesh &m = c->m;
    hipLaunchKernelGGL((kDp_node<S, TCH, CPB, SW, RS, SH>), dim3(xcd_grid((m.N2 + CPB - 1) / CPB), ny),
                       dim3(64 * (CPB + (SW ? 1 : 0))), 0, c->stream, m, c->d, c->t, dt, clamp, tpb);
}

template <bool SW>
static void launch_kdp_shape(fgpu_context *c, unsigned ny, double dt, int clamp, int tpb, int tch, int cpb)
{
    if (c->m.NL <= 64) {
        if (cpb == 8) launch_kdp<1, 2, 8, SW>(c, ny, dt, clamp, tpb);
        else if (cpb == 4 && tch == 4) launch_kdp<1, 4, 4, SW>(c, ny, dt, clamp, tpb);
        else if (cpb == 4) launch_kdp<1, 2, 4, SW>(c, ny, dt, clamp, tpb);
        else if (tch == 8) launch_kdp<1, 8, 1, SW>(c, ny, dt, clamp, tpb);
        else if (tch == 5) launch_kdp<1, 5, 1, SW>(c, ny, dt, clamp, tpb);
        else if (tch == 4) launch_kdp<1, 4, 1, SW>(c, ny, dt, clamp, tpb);
        else launch_kdp<1, 2, 1, SW>(c, ny, dt, clamp, tpb);
    } else if (c->m.NL <= KDP_SHORT_NL) {   // rows of KDP_SHORT_NL + 2
        // one column per block: R+- and the vertical fluxes by shuffles, which
        // frees their LDS rows (1.8 KB of a block's 10.1 KB at chunks of 5) for
        // more waves; 46.4-47.0 -> 45.7-46.0 ms at 1.3 M columns of 70 levels
        constexpr int R = KDP_SHORT_NL + 2;
        if (cpb >= 4) launch_kdp<2, 2, 4, SW, R>(c, ny, dt, clamp, tpb);
        else if (cpb == 3) launch_kdp<2, 2, 3, SW, R>(c, ny, dt, clamp, tpb);
        else if (cpb == 2 && tch == 4) launch_kdp<2, 4, 2, SW, R>(c, ny, dt, clamp, tpb);
        else if (cpb == 2) launch_kdp<2, 2, 2, SW, R>(c, ny, dt, clamp, tpb);
        else if (tch == 8) launch_kdp<2, 8, 1, SW, R, true>(c, ny, dt, clamp, tpb);
        else if (tch == 5) launch_kdp<2, 5, 1, SW, R, true>(c, ny, dt, clamp, tpb);
        else if (tch == 4) launch_kdp<2, 4, 1, SW, R, true>(c, ny, dt, clamp, tpb);
        else launch_kdp<2, 2, 1, SW, R, true>(c, ny, dt, clamp, tpb);
    } else {
        if (cpb >= 2 && tch == 4) launch_kdp<2, 4, 2, SW>(c, ny, dt, clamp, tpb);
        else if (cpb >= 2) launch_kdp<2, 2, 2, SW>(c, ny, dt, clamp, tpb);
        else if (tch == 8) launch_kdp<2, 8, 1, SW>(c, ny, dt, clamp, tpb);
        else if (tch == 5) launch_kdp<2, 5, 1, SW>(c, ny, dt, clamp, tpb);
        else if (tch == 4) launch_kdp<2, 4, 1, SW>(c, ny, dt, clamp, tpb);
        else launch_kdp<2, 2, 1, SW>(c, ny, dt, clamp, tpb);
    }
}

// K_C over (tracer, level) items with step_kcf tracers per block (1, 2 or 4)
template <int TPF, int NCH>
static void launch_kcf_t(fgpu_context *c, double dt)
{
    const DevMesh &m = c->m;
    const unsigned ny = (unsigned)((c->t.ntr + TPF - 1) / TPF);
    hipLaunchKernelGGL((kCf_node<TPF, NCH>), dim3(xcd_grid(m.N2), ny), dim3(64), 0, c->stream, m, c->t, dt);
}

static void launch_kcf(fgpu_context *c, double dt)
{
    const bool narrow = c->m.NL <= 64;
    if (c->step_kcf >= 4) {
        if (narrow) launch_kcf_t<4, 4>(c, dt);
        else launch_kcf_t<4, 8>(c, dt);
    } else if (c->step_kcf >= 2) {
        if (narrow) launch_kcf_t<2, 2>(c, dt);
        else launch_kcf_t<2, 4>(c, dt);
    } else {
        if (narrow) launch_kcf_t<1, 1>(c, dt);
        else launch_kcf_t<1, 2>(c, dt);
    }
}

int fgpu_enqueue_fused_step(fgpu_context *c, double dt)
{
    const DevMesh &m = c->m;
    const int nt = c->t.ntr;
    const int clamp = nt >= 2 ? 1 : -1;
    if (m.NL > 128) return fgpu_fail(FGPU_ERR_ARG, "more than 127 layers");
    // tracers per block: 1 runs every tracer's chain in its own waves
    const int tpb = std::max(1, std::min(c->step_tpb, nt));
    const unsigned ny = (unsigned)((nt + tpb - 1) / tpb);
    int tch, cpb;   // Thomas chunk and columns per block of K_D
    fgpu_kd_shape(c, &tch, &cpb);
    bool xv = false;   // some tracer uses CDIFF or PPM
    for (int v : c->h_ver) xv = xv || v == VER_CDIFF || v == VER_PPM;
    const bool ws = c->d.wsplit != 0;   // W split: implicit upwind LO step, w_i rows of the diffusion
    const int xf = fgpu_kd_extras(c);
    const bool ka2 = c->step_ka2 && m.L % 2 == 0;   // two levels per lane: 16-byte aligned rows only
    const bool kb2 = c->step_kb2 && m.L % 2 == 0 && !ws && !xv;
    {   // explicit diffusion terms from T^n (no-op without Ki, Redi or explicit vertical diffusion)
        int rc = fgpu_launch_diff(c, 0, nt, dt);
        if (rc) return rc;
    }
#define LAUNCH5(LBV)                                                                                                   \
    {                                                                                                                  \
        {                                                                                                              \
            ProfScope ps(c, "kG_node");                                                                                \
            hipLaunchKernelGGL(kG_node<LBV>, dim3(xcd_grid(m.N2), ny), dim3(LBV), 0, c->stream, m, c->t, c->d_kneed, tpb);      \
        }                                                                                                              \
        {                                                                                                              \
            ProfScope ps(c, "kA_edge");                                                                                \
            if (ka2)                                                                                                   \
                hipLaunchKernelGGL(kA2_edge, dim3(xcd_grid(m.E2), ny), dim3(64), 0, c->stream, m, c->d, c->t, tpb);    \
            else                                                                                                       \
                hipLaunchKernelGGL(kA_edge<LBV>, dim3(xcd_grid(m.E2), ny), dim3(LBV), 0, c->stream, m, c->d, c->t, tpb); \
        }                                                                                                              \
        {                                                                                                              \
            ProfScope ps(c, "kB_node");                                                                                \
            if (kb2)                                                                                                   \
                hipLaunchKernelGGL(kB2_node, dim3(xcd_grid(m.N2), ny), dim3(64), 0, c->stream, m, c->d, c->t, dt, tpb); \
            else if (ws)                                                                                               \
                hipLaunchKernelGGL((kB_node<LBV, true, true>), dim3(xcd_grid(m.N2), ny), dim3(LBV), 0, c->stream, m, c->d, c->t, dt, tpb); \
            else if (xv)                                                                                               \
                hipLaunchKernelGGL((kB_node<LBV, true, false>), dim3(xcd_grid(m.N2), ny), dim3(LBV), 0, c->stream, m, c->d, c->t, dt, tpb); \
            else                                                                                                       \
                hipLaunchKernelGGL((kB_node<LBV, false, false>), dim3(xcd_grid(m.N2), ny), dim3(LBV), 0, c->stream, m, c->d, c->t, dt, tpb); \
        }                                                                                                              \
        {                                                                                                              \
            ProfScope ps(c, "kC_node");                                                                                \
            if (c->step_kcf > 0)                                                                                       \
                launch_kcf(c, dt);                                                                                     \
            else                                                                                                       \
                hipLaunchKernelGGL(kC_node<LBV>, dim3(xcd_grid(m.N2), ny), dim3(LBV), 0, c->stream, m, c->t, dt, tpb); \
        }                                                                                                              \
        {                                                                                                              \
            ProfScope ps(c, "kD_node");                                                                                \
            if (xf == 0 && c->step_kdp == KD_SWEEPWAVE)                                                                \
                launch_kdp_shape<true>(c, ny, dt, clamp, tpb, tch, cpb);                                               \
            else if (xf == 0 && c->step_kdp == KD_WAVELOCAL)                                                           \
                launch_kdp_shape<false>(c, ny, dt, clamp, tpb, tch, cpb);                                              \
            else if (xf == 1)                                                                                          \
                hipLaunchKernelGGL((kD_node<LBV, 2, 1, 1>), dim3(xcd_grid(m.N2), ny), dim3(LBV), 0, c->stream, m, c->d, c->t, dt, clamp, tpb); \
            else if (xf == 2)                                                                                          \
                hipLaunchKernelGGL((kD_node<LBV, 2, 1, 2>), dim3(xcd_grid(m.N2), ny), dim3(LBV), 0, c->stream, m, c->d, c->t, dt, clamp, tpb); \
            else if (xf == 3)                                                                                          \
                hipLaunchKernelGGL((kD_node<LBV, 2, 1, 3>), dim3(xcd_grid(m.N2), ny), dim3(LBV), 0, c->stream, m, c->d, c->t, dt, clamp, tpb); \
            else if (LBV == 64 && cpb == 8)                                                                            \
                hipLaunchKernelGGL((kD_node<LBV, 2, (LBV == 64 ? 8 : 1), 0>), dim3(xcd_grid((m.N2 + 7) / 8), ny),          \
                                   dim3(LBV * (LBV == 64 ? 8 : 1)), 0, c->stream, m, c->d, c->t, dt, clamp, tpb);       \
            else if (LBV == 64 && cpb == 4 && tch == 4)                                                                \
                hipLaunchKernelGGL((kD_node<LBV, 4, (LBV == 64 ? 4 : 1), 0>), dim3(xcd_grid((m.N2 + 3) / 4), ny),          \
                                   dim3(LBV * (LBV == 64 ? 4 : 1)), 0, c->stream, m, c->d, c->t, dt, clamp, tpb);       \
            else if (LBV == 64 && cpb == 4)                                                                            \
                hipLaunchKernelGGL((kD_node<LBV, 2, (LBV == 64 ? 4 : 1), 0>), dim3(xcd_grid((m.N2 + 3) / 4), ny),          \
                                   dim3(LBV * (LBV == 64 ? 4 : 1)), 0, c->stream, m, c->d, c->t, dt, clamp, tpb);       \
            else if (tch == 8)                                                                                         \
                hipLaunchKernelGGL((kD_node<LBV, 8, 1, 0>), dim3(xcd_grid(m.N2), ny), dim3(LBV), 0, c->stream, m, c->d, c->t, dt, clamp, tpb); \
            else if (tch == 5)                                                                                         \
                hipLaunchKernelGGL((kD_node<LBV, 5, 1, 0>), dim3(xcd_grid(m.N2), ny), dim3(LBV), 0, c->stream, m, c->d, c->t, dt, clamp, tpb); \
            else if (tch == 4)                                                                                         \
                hipLaunchKernelGGL((kD_node<LBV, 4, 1, 0>), dim3(xcd_grid(m.N2), ny), dim3(LBV), 0, c->stream, m, c->d, c->t, dt, clamp, tpb); \
            else                                                                                                       \
                hipLaunchKernelGGL((kD_node<LBV, 2, 1, 0>), dim3(xcd_grid(m.N2), ny), dim3(LBV), 0, c->stream, m, c->d, c->t, dt, clamp, tpb); \
        }                                                                                                              \
    }
    // one lane per level and per interface: LB >= NL
    if (m.NL <= 64) LAUNCH5(64)
    else LAUNCH5(128)
#undef LAUNCH5
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fgpu_fail(FGPU_ERR_HIP, std::string("fused step launch: ") + hipGetErrorString(e));
    return FGPU_OK;
}
