/*
 * fesom_oracle.c -- CPU oracle (TEST INFRASTRUCTURE, never the product path).
 *
 * A single-threaded C restatement of the reference's tracer-transport step and
 * SSH solve. Each routine names the reference file:line it follows. Loops run
 * in the reference's serial order (edges ascending, elements in nod_in_elem2D
 * order), so the edge->node scatters sum in the order a single-threaded
 * reference run uses. Floating-point expressions keep the reference's
 * evaluation order; build with -ffp-contract=off.
 *
 * fo_set_mutation(m) plants one deliberate mistake (m != 0) so the pinning
 * tests can show that they catch it:
 *   1  AB2 coefficient drops epsilon                      (init_tracers_AB)
 *   2  y-gradient uses the x coefficients                 (tracer_gradient_elements)
 *   3  upwind-triangle search does not negate the edge    (find_up_downwind_triangles)
 *   4  shared levels take the upwind gradient for both    (fill_up_dn_grad)
 *   5  second-element volume flux with the wrong sign     (adv_tra_hor_upw1)
 *   6  MFCT reconstruction weight 3 instead of 2          (adv_tra_hor_mfct)
 *   7  QR4C swaps the upper and lower slopes              (adv_tra_ver_qr4c)
 *   8  FCT vertical limiter uses R+ of the upper cell for an upward flux
 *   10 flux2dtracer divides node 2 by node 1's volume     (oce_tra_adv_flux2dtracer)
 *   12 implicit diffusion uses area(nz) in c(nz)          (diff_ver_part_impl_ale)
 *   13 Thomas back-substitution adds instead of subtracts
 *   14 SSH diagonal adds areasvol*dt instead of /dt       (init_stiff_mat_ale)
 *   15 CG beta inverted                                   (ssh_solve_cg)
 *   16 preconditioner variant 1 off-diagonal sign flipped (ssh_solve_preconditioner)
 *   17 upw1 vertical flux uses T(k) on both sides         (adv_tra_ver_upw1)
 *   18 FCT b1 drops the vertical contribution to P+/P-    (oce_tra_adv_fct)
 *   19 second-element transport added with the first's sign (vert_vel_ale)
 *   20 W accumulated from the top instead of the bottom   (vert_vel_ale)
 *   21 SSH RHS drops helem in the first element's term    (compute_ssh_rhs_ale)
 *   22 CFL_z below an interface uses W of the interface above (compute_CFLz)
 *   23 stiffness update adds +dhe instead of -dhe        (update_stiff_mat_ale)
 *   24 PPM upward flux takes the parabola's lower edge    (adv_tra_vert_ppm)
 *   25 CDIFF interface value is the lower cell's, not the mean (adv_tra_ver_cdiff)
 *   26 nboundary_lay not reset at the nodes of boundary edges (muscl_adv_init)
 *   27 implicit advection's lower diagonal takes max(0, w) (adv_tra_vert_impl)
 *   28 surface w_i term of the implicit diffusion with the wrong sign (diff_ver_part_impl_ale)
 *   29 explicit vertical diffusion divides by the interface spacing above (diff_ver_part_expl_ale)
 *   30 horizontal diffusion, two-sided levels: second element's dual segment with the wrong sign
 *      (diff_part_hor_redi)
 *   31 horizontal diffusion takes Ki of the first node only (diff_part_hor_redi)
 *   32 Redi vertical flux weights the upper layer by the distance to the interface below
 *      (diff_ver_part_redi_expl)
 *   33 Redi K33 below an interior interface scaled by the spacing below (diff_ver_part_impl_ale)
 *   34 hbar updated without dividing by the node's area (compute_hbar_ale)
 *   35 dhe averages two of the element's three nodes (compute_hbar_ale)
 *   36 zlevel corrects W at a layer's top by that layer's change alone (vert_vel_ale)
 *   37 zlevel thins the top layers down to zero instead of min_hnode (vert_vel_ale)
 *   38 SSH diagonal gets the mass term under an ice-shelf cavity too (init_stiff_mat_ale)
 *   39 zstar stretches the cavity columns as well                (vert_vel_ale)
 *   40 ssh_rhs_old enters the full free surface RHS at cavity nodes (compute_ssh_rhs_ale)
 *   41 surface flux added to layer 1 instead of the first wet layer (diff_ver_part_impl_ale)
 *   42 UPW1 skips the top segments (A, B) where only one element is wet (adv_tra_hor_upw1)
 *   43 MFCT skips the top segment (B) where only the second element is wet (adv_tra_hor_mfct)
 *   44 UPW1 bottom segment (E) takes the first element's velocity and thickness (adv_tra_hor_upw1)
 *   45 AB3 swaps the weights of the two old levels                 (init_tracers_AB)
 *   46 ice-shelf meltwater (ID 501) enters open-ocean columns too   (bc_surface)
 *   47 the isotope IDs 102 and 104 take each other's flux column   (bc_surface)
 *   48 depth-mode hosing (ID 304) reads the layer below nzmin     (bc_surface)
 *   49 the node bounds leave out elements without the level (no +-1e3 padding) (oce_tra_adv_fct)
 *   50 QR4C's second interface takes the lower layer's value (upwind) instead of the mean (adv_tra_ver_qr4c)
 *   51 QR4C applies its coinciding second / last-but-one row once in two-layer columns (adv_tra_ver_qr4c)
 *   52 the node-mean gradient of fill_up_dn_grad is an unweighted mean over the elements (fill_up_dn_grad)
 *   53 the FCT denominators leave out flux_eps                       (oce_tra_adv_fct)
 */
#include "fesom_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

/* The OpenMP build (libfesom_oracle_omp.so, the multi-core CPU baseline) runs
 * the column, edge and element loops in parallel and replaces each edge->node
 * scatter by a gather over the node's edges in ascending edge order, which
 * adds the same terms in the same order: it is bit-identical to the serial
 * build, which a test checks. The planted mistakes are for the serial build. */
#ifdef _OPENMP
#define OMP_FOR _Pragma("omp parallel for schedule(static)")
#else
#define OMP_FOR
#endif

static int g_mut = 0;
void fo_set_mutation(int m) { g_mut = m; }
int fo_get_mutation(void) { return g_mut; }

#define R_EARTH 6367500.0
#define G_ACC 9.81
#define EPS_AB 0.1
#define VCPW 4.2e6

#define LAY(n, k) ((size_t)(n) * (size_t)L + (size_t)((k) - 1))
#define INT(n, k) ((size_t)(n) * (size_t)NL + (size_t)((k) - 1))

static inline double dmax(double a, double b) { return a > b ? a : b; }
static inline double dmin(double a, double b) { return a < b ? a : b; }
static inline int imax(int a, int b) { return a > b ? a : b; }
static inline int imin(int a, int b) { return a < b ? a : b; }

/* ------------------------------------------------------------------------ */
/* find_up_downwind_triangles: oce_muscl_adv.F90:162-352                     */
/* ------------------------------------------------------------------------ */
static double trim_cyc(double b, double Lc)
{   /* gen_modules_rotate_grid.F90:204-211 */
    if (b > Lc / 2.0) b = b - Lc;
    if (b < -Lc / 2.0) b = b + Lc;
    return b;
}

static int updn_search(const fo_mesh *m, int node, const double x[2])
{
    const double Lc = m->cyclic_length;
    int found = -1;
    for (int k = 0; k < m->nod_in_elem2D_num[node]; ++k) {
        int elem = m->nod_in_elem2D[(size_t)node * m->kmax + k];
        const int *en = m->elem2D_nodes + 3 * (size_t)elem;
        int a0, a1, a2;
        if (en[0] == node) { a0 = 0; a1 = 1; a2 = 2; }
        else if (en[1] == node) { a0 = 1; a1 = 0; a2 = 2; }
        else { a0 = 2; a1 = 0; a2 = 1; }
        double b[2], c[2];
        b[0] = m->coord_nod2D[2 * en[a1]] - m->coord_nod2D[2 * en[a0]];
        b[1] = m->coord_nod2D[2 * en[a1] + 1] - m->coord_nod2D[2 * en[a0] + 1];
        c[0] = m->coord_nod2D[2 * en[a2]] - m->coord_nod2D[2 * en[a0]];
        c[1] = m->coord_nod2D[2 * en[a2] + 1] - m->coord_nod2D[2 * en[a0] + 1];
        b[0] = trim_cyc(b[0], Lc);
        c[0] = trim_cyc(c[0], Lc);
        double cr = c[0] * c[0] + c[1] * c[1];
        double bx = (b[0] * c[0] + b[1] * c[1]) / cr;
        double by = (-b[0] * c[1] + b[1] * c[0]) / cr;
        double xx = (x[0] * c[0] + x[1] * c[1]) / cr;
        double xy = (-x[0] * c[1] + x[1] * c[0]) / cr;
        double ab = atan2(by, bx);
        double ax = atan2(xy, xx);
        /* every branch is a Fortran CYCLE: the last matching element wins */
        if ((ab > 0.0) && (ax > 0.0) && (ax < ab)) { found = elem; continue; }
        if ((ab < 0.0) && (ax < 0.0) && (ax > ab)) { found = elem; continue; }
        if ((ab == ax) || (ax == 0.0)) { found = elem; continue; }
    }
    return found;
}

int fo_find_up_downwind_triangles(const fo_mesh *m, int *up_dn)
{
    const double Lc = m->cyclic_length;
    for (int e = 0; e < m->edge2D; ++e) {
        int n1 = m->edges[2 * e], n2 = m->edges[2 * e + 1];
        double x[2];
        x[0] = m->coord_nod2D[2 * n2] - m->coord_nod2D[2 * n1];
        x[1] = m->coord_nod2D[2 * n2 + 1] - m->coord_nod2D[2 * n1 + 1];
        if (x[0] > Lc / 2.0) x[0] = x[0] - Lc;
        if (x[0] < -Lc / 2.0) x[0] = x[0] + Lc;
        double xm[2] = {-x[0], -x[1]};
        if (g_mut == 3) { xm[0] = x[0]; xm[1] = x[1]; }
        up_dn[2 * e] = updn_search(m, n1, xm);
        up_dn[2 * e + 1] = updn_search(m, n2, x);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* tracer_gradient_elements: oce_tracer_mod.F90:147-188                      */
/* ------------------------------------------------------------------------ */
void fo_tracer_gradient_elements(const fo_mesh *m, const double *ttf, double *tr_xy)
{
    const int L = m->nl - 1;
    OMP_FOR
    for (int el = 0; el < m->elem2D; ++el) {
        const int *en = m->elem2D_nodes + 3 * (size_t)el;
        const double *gs = m->gradient_sca + 6 * (size_t)el;
        for (int nz = m->ulevels[el]; nz <= m->nlevels[el] - 1; ++nz) {
            double t0 = ttf[LAY(en[0], nz)], t1 = ttf[LAY(en[1], nz)], t2 = ttf[LAY(en[2], nz)];
            double gx = gs[0] * t0 + gs[1] * t1 + gs[2] * t2;
            double gy = gs[3] * t0 + gs[4] * t1 + gs[5] * t2;
            if (g_mut == 2) gy = gs[0] * t0 + gs[1] * t1 + gs[2] * t2;
            tr_xy[LAY(el, nz) * 2 + 0] = gx;
            tr_xy[LAY(el, nz) * 2 + 1] = gy;
        }
    }
}

/* ------------------------------------------------------------------------ */
/* fill_up_dn_grad: oce_muscl_adv.F90:356-525                                */
/* ------------------------------------------------------------------------ */
static void node_mean_grad(const fo_mesh *m, const double *tr_xy, int node, int nz,
                           double *gx, double *gy)
{
    const int L = m->nl - 1;
    double tvol = 0.0, tx = 0.0, ty = 0.0;
    for (int k = 0; k < m->nod_in_elem2D_num[node]; ++k) {
        int elem = m->nod_in_elem2D[(size_t)node * m->kmax + k];
        if (m->nlevels[elem] - 1 < nz || nz < m->ulevels[elem]) continue;
        const double wgt = (g_mut == 52) ? 1.0 : m->elem_area[elem];
        tvol = tvol + wgt;
        tx = tx + tr_xy[LAY(elem, nz) * 2 + 0] * wgt;
        ty = ty + tr_xy[LAY(elem, nz) * 2 + 1] * wgt;
    }
    *gx = tx / tvol;
    *gy = ty / tvol;
}

void fo_fill_up_dn_grad(const fo_mesh *m, const double *tr_xy, double *grad)
{
    const int L = m->nl - 1;
    OMP_FOR
    for (int e = 0; e < m->edge2D; ++e) {
        int n1 = m->edges[2 * e], n2 = m->edges[2 * e + 1];
        int up = m->edge_up_dn_tri[2 * e], dn = m->edge_up_dn_tri[2 * e + 1];
        double gx, gy;
        if (up >= 0 && dn >= 0) {
            int nzmin = imax(m->ulevels_nod2D_max[n1], m->ulevels_nod2D_max[n2]);
            int nzmax = imin(m->nlevels_nod2D_min[n1], m->nlevels_nod2D_min[n2]);
            for (int nz = m->ulevels_nod2D[n1]; nz <= nzmin - 1; ++nz) {
                node_mean_grad(m, tr_xy, n1, nz, &gx, &gy);
                grad[LAY(e, nz) * 4 + 0] = gx;
                grad[LAY(e, nz) * 4 + 2] = gy;
            }
            for (int nz = m->ulevels_nod2D[n2]; nz <= nzmin - 1; ++nz) {
                node_mean_grad(m, tr_xy, n2, nz, &gx, &gy);
                grad[LAY(e, nz) * 4 + 1] = gx;
                grad[LAY(e, nz) * 4 + 3] = gy;
            }
            for (int nz = nzmin; nz <= nzmax - 1; ++nz) {
                int dn_use = (g_mut == 4) ? up : dn;
                grad[LAY(e, nz) * 4 + 0] = tr_xy[LAY(up, nz) * 2 + 0];
                grad[LAY(e, nz) * 4 + 1] = tr_xy[LAY(dn_use, nz) * 2 + 0];
                grad[LAY(e, nz) * 4 + 2] = tr_xy[LAY(up, nz) * 2 + 1];
                grad[LAY(e, nz) * 4 + 3] = tr_xy[LAY(dn_use, nz) * 2 + 1];
            }
            for (int nz = nzmax; nz <= m->nlevels_nod2D[n1] - 1; ++nz) {
                node_mean_grad(m, tr_xy, n1, nz, &gx, &gy);
                grad[LAY(e, nz) * 4 + 0] = gx;
                grad[LAY(e, nz) * 4 + 2] = gy;
            }
            for (int nz = nzmax; nz <= m->nlevels_nod2D[n2] - 1; ++nz) {
                node_mean_grad(m, tr_xy, n2, nz, &gx, &gy);
                grad[LAY(e, nz) * 4 + 1] = gx;
                grad[LAY(e, nz) * 4 + 3] = gy;
            }
        } else {
            for (int nz = m->ulevels_nod2D[n1]; nz <= m->nlevels_nod2D[n1] - 1; ++nz) {
                node_mean_grad(m, tr_xy, n1, nz, &gx, &gy);
                grad[LAY(e, nz) * 4 + 0] = gx;
                grad[LAY(e, nz) * 4 + 2] = gy;
            }
            for (int nz = m->ulevels_nod2D[n2]; nz <= m->nlevels_nod2D[n2] - 1; ++nz) {
                node_mean_grad(m, tr_xy, n2, nz, &gx, &gy);
                grad[LAY(e, nz) * 4 + 1] = gx;
                grad[LAY(e, nz) * 4 + 3] = gy;
            }
        }
    }
}

/* ------------------------------------------------------------------------ */
/* init_tracers_AB: oce_tracer_mod.F90:13-143                                */
/* ------------------------------------------------------------------------ */
int fo_init_tracers_AB(const fo_mesh *m, fo_tracer *t, fo_work *w)
{
    const int L = m->nl - 1;
    const size_t n3 = (size_t)m->nod2D * L;
    memset(w->del_ttf, 0, n3 * sizeof(double));
    memset(w->del_ttf_advhoriz, 0, n3 * sizeof(double));
    memset(w->del_ttf_advvert, 0, n3 * sizeof(double));
    if (t->AB_order == 2) {
        double c1 = (g_mut == 1) ? 1.5 : (1.5 + EPS_AB);
        OMP_FOR
        for (size_t i = 0; i < n3; ++i) {
            t->valuesAB[i] = -(0.5 + EPS_AB) * t->valuesold[i] + c1 * t->values[i];
            t->valuesold[i] = t->values[i];
        }
    } else if (t->AB_order == 3) {
        for (size_t i = 0; i < n3; ++i) {
            double v = (g_mut == 45) ? 5.0 * t->valuesold[2 * i] - 16.0 * t->valuesold[2 * i + 1] + 23.0 * t->values[i]
                                     : 5.0 * t->valuesold[2 * i + 1] - 16.0 * t->valuesold[2 * i] + 23.0 * t->values[i];
            t->valuesAB[i] = v / 12.0;
        }
        for (size_t i = 0; i < n3; ++i) {
            t->valuesold[2 * i + 1] = t->valuesold[2 * i];
            t->valuesold[2 * i] = t->values[i];
        }
    } else {
        return -1;   /* par_ex: AB order must be 2 or 3 (oce_tracer_mod.F90:75-91) */
    }
    /* gradients from values, not valuesAB (oce_tracer_mod.F90:126-131) */
    fo_tracer_gradient_elements(m, t->values, w->tr_xy);
    if (w->tr_z) fo_tracer_gradient_z(m, t->values, w->tr_z);
    fo_fill_up_dn_grad(m, w->tr_xy, w->edge_up_dn_grad);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* tracer_gradient_z: oce_tracer_mod.F90:192-232 (old thicknesses hnode)     */
/* ------------------------------------------------------------------------ */
void fo_tracer_gradient_z(const fo_mesh *m, const double *ttf, double *tr_z)
{
    const int L = m->nl - 1, NL = m->nl;
    for (int n = 0; n < m->nod2D; ++n) {
        const int nzmax = m->nlevels_nod2D[n], nzmin = m->ulevels_nod2D[n];
        for (int nz = nzmin + 1; nz <= nzmax - 1; ++nz) {
            const double dz = 0.5 * (m->hnode[LAY(n, nz - 1)] + m->hnode[LAY(n, nz)]);
            tr_z[INT(n, nz)] = (ttf[LAY(n, nz - 1)] - ttf[LAY(n, nz)]) / dz;
        }
        tr_z[INT(n, nzmin)] = 0.0;
        tr_z[INT(n, nzmax)] = 0.0;
    }
    (void)L;
}

/* ------------------------------------------------------------------------ */
/* adv_tra_hor_upw1: oce_adv_tra_hor.F90:64-257                              */
/* ------------------------------------------------------------------------ */
typedef struct { int n1, n2, el1, el2, nl1, nu1, nl2, nu2; } edge_info;

static edge_info get_edge(const fo_mesh *m, int e)
{
    edge_info q;
    q.n1 = m->edges[2 * e];
    q.n2 = m->edges[2 * e + 1];
    q.el1 = m->edge_tri[2 * e];
    q.el2 = m->edge_tri[2 * e + 1];
    q.nl1 = m->nlevels[q.el1] - 1;
    q.nu1 = m->ulevels[q.el1];
    q.nl2 = 0;
    q.nu2 = 0;
    if (q.el2 >= 0) {
        q.nl2 = m->nlevels[q.el2] - 1;
        q.nu2 = m->ulevels[q.el2];
    }
    return q;
}

static inline double upw1_flux(double t1, double t2, double vflux, double fin)
{
    return -0.5 * (t1 * (vflux + fabs(vflux)) + t2 * (vflux - fabs(vflux))) - fin;
}

void fo_adv_tra_hor_upw1(const fo_mesh *m, const double *vel, const double *ttf, double *flux, int init_zero)
{
    const int L = m->nl - 1;
    if (init_zero) memset(flux, 0, (size_t)m->edge2D * L * sizeof(double));
    OMP_FOR
    for (int e = 0; e < m->edge2D; ++e) {
        edge_info q = get_edge(m, e);
        double dX1 = m->edge_cross_dxdy[4 * e + 0], dY1 = m->edge_cross_dxdy[4 * e + 1];
        double dX2 = 0.0, dY2 = 0.0;
        if (q.el2 >= 0) { dX2 = m->edge_cross_dxdy[4 * e + 2]; dY2 = m->edge_cross_dxdy[4 * e + 3]; }
        int nl12 = imin(q.nl1, q.nl2);
        int nu12 = imax(q.nu1, q.nu2);
        const double s2 = (g_mut == 5) ? -1.0 : 1.0;
#define V1(c, nz) vel[(LAY(q.el1, nz)) * 2 + (c)]
#define V2(c, nz) vel[(LAY(q.el2, nz)) * 2 + (c)]
#define H1(nz) m->helem[LAY(q.el1, nz)]
#define H2(nz) m->helem[LAY(q.el2, nz)]
        for (int nz = q.nu1; nz <= nu12 - 1 && g_mut != 42; ++nz) {   /* (A) */
            double vflux = (-V1(1, nz) * dX1 + V1(0, nz) * dY1) * H1(nz);
            flux[LAY(e, nz)] = upw1_flux(ttf[LAY(q.n1, nz)], ttf[LAY(q.n2, nz)], vflux, flux[LAY(e, nz)]);
        }
        if (q.nu2 > 0 && g_mut != 42) {
            for (int nz = q.nu2; nz <= nu12 - 1; ++nz) {            /* (B) */
                double vflux = s2 * (V2(1, nz) * dX2 - V2(0, nz) * dY2) * H2(nz);
                flux[LAY(e, nz)] = upw1_flux(ttf[LAY(q.n1, nz)], ttf[LAY(q.n2, nz)], vflux, flux[LAY(e, nz)]);
            }
        }
        for (int nz = nu12; nz <= nl12; ++nz) {                    /* (C) */
            double vflux = (-V1(1, nz) * dX1 + V1(0, nz) * dY1) * H1(nz)
                         + s2 * (V2(1, nz) * dX2 - V2(0, nz) * dY2) * H2(nz);
            flux[LAY(e, nz)] = upw1_flux(ttf[LAY(q.n1, nz)], ttf[LAY(q.n2, nz)], vflux, flux[LAY(e, nz)]);
        }
        for (int nz = nl12 + 1; nz <= q.nl1; ++nz) {               /* (D) */
            double vflux = (-V1(1, nz) * dX1 + V1(0, nz) * dY1) * H1(nz);
            flux[LAY(e, nz)] = upw1_flux(ttf[LAY(q.n1, nz)], ttf[LAY(q.n2, nz)], vflux, flux[LAY(e, nz)]);
        }
        for (int nz = nl12 + 1; nz <= q.nl2; ++nz) {               /* (E) */
            double vflux = (g_mut == 44) ? (V1(1, nz) * dX2 - V1(0, nz) * dY2) * H1(nz)
                                         : s2 * (V2(1, nz) * dX2 - V2(0, nz) * dY2) * H2(nz);
            flux[LAY(e, nz)] = upw1_flux(ttf[LAY(q.n1, nz)], ttf[LAY(q.n2, nz)], vflux, flux[LAY(e, nz)]);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* adv_tra_hor_mfct: oce_adv_tra_hor.F90:546-834                             */
/* ------------------------------------------------------------------------ */
/* MFCT (nblay NULL) and MUSCL: MUSCL scales each node's reconstruction term by
   c_lo = 1 at levels down to nboundary_lay(node) and 0 below
   (oce_adv_tra_hor.F90:261-542, c_lo at :356-357) */
static void hor_mfct_muscl(const fo_mesh *m, const double *vel, const double *ttf, double num_ord, double *flux,
                           const double *grad, const int *nblay, int init_zero)
{
    const int L = m->nl - 1;
    const double w2 = (g_mut == 6) ? 3.0 : 2.0;
    if (init_zero) memset(flux, 0, (size_t)m->edge2D * L * sizeof(double));
    OMP_FOR
    for (int e = 0; e < m->edge2D; ++e) {
        edge_info q = get_edge(m, e);
        double dX1 = m->edge_cross_dxdy[4 * e + 0], dY1 = m->edge_cross_dxdy[4 * e + 1];
        double dX2 = 0.0, dY2 = 0.0;
        double a = R_EARTH * m->elem_cos[q.el1];
        if (q.el2 >= 0) {
            dX2 = m->edge_cross_dxdy[4 * e + 2];
            dY2 = m->edge_cross_dxdy[4 * e + 3];
            a = 0.5 * (a + R_EARTH * m->elem_cos[q.el2]);
        }
        const double edx = m->edge_dxdy[2 * e], edy = m->edge_dxdy[2 * e + 1];
        int nl12 = imin(q.nl1, q.nl2);
        int nu12 = imax(q.nu1, q.nu2);
        for (int pass = 0; pass < 5; ++pass) {
            int lo, hi;
            if (pass == 0) { lo = q.nu1; hi = nu12 - 1; }
            else if (pass == 1) { if (q.nu2 <= 0 || g_mut == 43) continue; lo = q.nu2; hi = nu12 - 1; }
            else if (pass == 2) { lo = nu12; hi = nl12; }
            else if (pass == 3) { lo = nl12 + 1; hi = q.nl1; }
            else { lo = nl12 + 1; hi = q.nl2; }
            for (int nz = lo; nz <= hi; ++nz) {
                const double T1 = ttf[LAY(q.n1, nz)], T2 = ttf[LAY(q.n2, nz)];
                const double *g = grad + LAY(e, nz) * 4;
                double Tmean2, Tmean1;
                if (nblay) {
                    const double c1 = (double)imax((nblay[q.n1] - nz) >= 0 ? 1 : -1, 0);
                    const double c2 = (double)imax((nblay[q.n2] - nz) >= 0 ? 1 : -1, 0);
                    Tmean2 = T2 - (w2 * (T2 - T1) + edx * a * g[1] + edy * R_EARTH * g[3]) / 6.0 * c2;
                    Tmean1 = T1 + (w2 * (T2 - T1) + edx * a * g[0] + edy * R_EARTH * g[2]) / 6.0 * c1;
                } else {
                    Tmean2 = T2 - (w2 * (T2 - T1) + edx * a * g[1] + edy * R_EARTH * g[3]) / 6.0;
                    Tmean1 = T1 + (w2 * (T2 - T1) + edx * a * g[0] + edy * R_EARTH * g[2]) / 6.0;
                }
                double vflux;
                if (pass == 0 || pass == 3)
                    vflux = (-V1(1, nz) * dX1 + V1(0, nz) * dY1) * H1(nz);
                else if (pass == 1 || pass == 4)
                    vflux = (V2(1, nz) * dX2 - V2(0, nz) * dY2) * H2(nz);
                else
                    vflux = (-V1(1, nz) * dX1 + V1(0, nz) * dY1) * H1(nz)
                          + (V2(1, nz) * dX2 - V2(0, nz) * dY2) * H2(nz);
                double cHO = (vflux + fabs(vflux)) * Tmean1 + (vflux - fabs(vflux)) * Tmean2;
                flux[LAY(e, nz)] = -0.5 * (1.0 - num_ord) * cHO
                                   - vflux * num_ord * 0.5 * (Tmean1 + Tmean2) - flux[LAY(e, nz)];
            }
        }
    }
#undef V1
#undef V2
#undef H1
#undef H2
}

void fo_adv_tra_hor_mfct(const fo_mesh *m, const double *vel, const double *ttf, double num_ord,
                         double *flux, const double *grad, int init_zero)
{
    hor_mfct_muscl(m, vel, ttf, num_ord, flux, grad, NULL, init_zero);
}

void fo_adv_tra_hor_muscl(const fo_mesh *m, const double *vel, const double *ttf, double num_ord,
                          double *flux, const double *grad, const int *nblay, int init_zero)
{
    hor_mfct_muscl(m, vel, ttf, num_ord, flux, grad, nblay, init_zero);
}

/* nboundary_lay of muscl_adv_init (oce_muscl_adv.F90:91-157): nl-1, reset to 0
   at both nodes of a boundary edge, else the minimum over the node's edges of
   min(nlevels of the two triangles) - 1, edges in order */
void fo_muscl_nboundary_lay(const fo_mesh *m, int *nblay)
{
    for (int n = 0; n < m->nod2D; ++n) nblay[n] = m->nl - 1;
    for (int e = 0; e < m->edge2D; ++e) {
        const int n1 = m->edges[2 * e], n2 = m->edges[2 * e + 1];
        const int el1 = m->edge_tri[2 * e], el2 = m->edge_tri[2 * e + 1];
        if (el1 < 0 || el2 < 0) {
            if (g_mut != 26) { nblay[n1] = 0; nblay[n2] = 0; }
        } else {
            const int v = imin(m->nlevels[el1], m->nlevels[el2]) - 1;
            nblay[n1] = imin(nblay[n1], v);
            nblay[n2] = imin(nblay[n2], v);
        }
    }
}

/* ------------------------------------------------------------------------ */
/* adv_tra_ver_upw1: oce_adv_tra_ver.F90:244-328                             */
/* ------------------------------------------------------------------------ */
void fo_adv_tra_ver_upw1(const fo_mesh *m, const double *W, const double *ttf, double *flux, int init_zero)
{
    const int L = m->nl - 1, NL = m->nl;
    if (init_zero) memset(flux, 0, (size_t)m->nod2D * NL * sizeof(double));
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nzmax = m->nlevels_nod2D[n], nzmin = m->ulevels_nod2D[n];
        int nz = nzmin;
        flux[INT(n, nz)] = -W[INT(n, nz)] * ttf[LAY(n, nz)] * m->area[INT(n, nz)] - flux[INT(n, nz)];
        nz = nzmax;
        flux[INT(n, nz)] = 0.0 - flux[INT(n, nz)];
        for (nz = nzmin + 1; nz <= nzmax - 1; ++nz) {
            double w = W[INT(n, nz)];
            double tup = (g_mut == 17) ? ttf[LAY(n, nz)] : ttf[LAY(n, nz - 1)];
            flux[INT(n, nz)] = -0.5 * (ttf[LAY(n, nz)] * (w + fabs(w)) + tup * (w - fabs(w)))
                               * m->area[INT(n, nz)] - flux[INT(n, nz)];
        }
    }
}

/* ------------------------------------------------------------------------ */
/* adv_tra_ver_qr4c: oce_adv_tra_ver.F90:332-434                             */
/* ------------------------------------------------------------------------ */
void fo_adv_tra_ver_qr4c(const fo_mesh *m, const double *W, const double *ttf, double num_ord,
                         double *flux, int init_zero)
{
    const int L = m->nl - 1, NL = m->nl;
    if (init_zero) memset(flux, 0, (size_t)m->nod2D * NL * sizeof(double));
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nzmax = m->nlevels_nod2D[n], nzmin = m->ulevels_nod2D[n];
        const double *T = ttf + LAY(n, 1);
        const double *Zn = m->Z_3d_n + LAY(n, 1);
        const double *zb = m->zbar_3d_n + INT(n, 1);
        const double *w = W + INT(n, 1);
        const double *ar = m->area + INT(n, 1);
        double *f = flux + INT(n, 1);
        /* 1-based level k lives at index k-1 */
#define I(k) ((k) - 1)
        int nz = nzmin;
        f[I(nz)] = -T[I(nz)] * w[I(nz)] * ar[I(nz)] - f[I(nz)];
        nz = nzmin + 1;
        if (g_mut == 50) f[I(nz)] = -T[I(nz)] * w[I(nz)] * ar[I(nz)] - f[I(nz)];
        else f[I(nz)] = -0.5 * (T[I(nz - 1)] + T[I(nz)]) * w[I(nz)] * ar[I(nz)] - f[I(nz)];
        nz = nzmax - 1;
        /* with two layers nzmax-1 = nzmin+1: the reference applies this row a
           second time to the flux just written (oce_adv_tra_ver.F90:392-400) */
        if (!(g_mut == 51 && nzmax - 1 == nzmin + 1))
            f[I(nz)] = -0.5 * (T[I(nz - 1)] + T[I(nz)]) * w[I(nz)] * ar[I(nz)] - f[I(nz)];
        nz = nzmax;
        f[I(nz)] = 0.0 - f[I(nz)];
        for (nz = nzmin + 2; nz <= nzmax - 2; ++nz) {
            double qc = (T[I(nz - 1)] - T[I(nz)]) / (Zn[I(nz - 1)] - Zn[I(nz)]);
            double qu = (T[I(nz)] - T[I(nz + 1)]) / (Zn[I(nz)] - Zn[I(nz + 1)]);
            double qd = (T[I(nz - 2)] - T[I(nz - 1)]) / (Zn[I(nz - 2)] - Zn[I(nz - 1)]);
            if (g_mut == 7) { double s = qu; qu = qd; qd = s; }
            double Tmean1 = T[I(nz)] + (2 * qc + qu) * (zb[I(nz)] - Zn[I(nz)]) / 3.0;
            double Tmean2 = T[I(nz - 1)] + (2 * qc + qd) * (zb[I(nz)] - Zn[I(nz - 1)]) / 3.0;
            double ww = w[I(nz)];
            double Tmean = (ww + fabs(ww)) * Tmean1 + (ww - fabs(ww)) * Tmean2;
            f[I(nz)] = (-0.5 * (1.0 - num_ord) * Tmean - num_ord * (0.5 * (Tmean1 + Tmean2)) * ww)
                       * ar[I(nz)] - f[I(nz)];
        }
#undef I
    }
}

/* ------------------------------------------------------------------------ */
/* oce_tra_adv_fct: oce_adv_tra_fct.F90:72-512                               */
/* ------------------------------------------------------------------------ */
void fo_oce_tra_adv_fct(const fo_mesh *m, double dt, const double *ttf, const double *lo,
                        double *adf_h, double *adf_v, double *fct_ttf_min, double *fct_ttf_max,
                        double *fct_plus, double *fct_minus)
{
    const int L = m->nl - 1, NL = m->nl;
    const double flux_eps = (g_mut == 53) ? 0.0 : 1e-16, bignumber = 1e3;
    double *aux = (double *)calloc((size_t)m->elem2D * L * 2, sizeof(double));
    double *tvmax = (double *)calloc((size_t)m->nod2D * L, sizeof(double));
    double *tvmin = (double *)calloc((size_t)m->nod2D * L, sizeof(double));
    /* a1 */
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n)
        for (int nz = m->ulevels_nod2D[n]; nz <= m->nlevels_nod2D[n] - 1; ++nz) {
            fct_ttf_max[LAY(n, nz)] = dmax(lo[LAY(n, nz)], ttf[LAY(n, nz)]);
            fct_ttf_min[LAY(n, nz)] = dmin(lo[LAY(n, nz)], ttf[LAY(n, nz)]);
        }
    /* a2 */
    OMP_FOR
    for (int el = 0; el < m->elem2D; ++el) {
        const int *en = m->elem2D_nodes + 3 * (size_t)el;
        int nu1 = m->ulevels[el], nl1 = m->nlevels[el];
        for (int nz = 1; nz <= nu1 - 1; ++nz) {
            aux[LAY(el, nz) * 2 + 0] = -bignumber;
            aux[LAY(el, nz) * 2 + 1] = bignumber;
        }
        for (int nz = nu1; nz <= nl1 - 1; ++nz) {
            aux[LAY(el, nz) * 2 + 0] = dmax(dmax(fct_ttf_max[LAY(en[0], nz)], fct_ttf_max[LAY(en[1], nz)]),
                                            fct_ttf_max[LAY(en[2], nz)]);
            aux[LAY(el, nz) * 2 + 1] = dmin(dmin(fct_ttf_min[LAY(en[0], nz)], fct_ttf_min[LAY(en[1], nz)]),
                                            fct_ttf_min[LAY(en[2], nz)]);
        }
        for (int nz = nl1; nz <= L; ++nz) {
            aux[LAY(el, nz) * 2 + 0] = -bignumber;
            aux[LAY(el, nz) * 2 + 1] = bignumber;
        }
    }
    /* a3 */
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        const int *ne = m->nod_in_elem2D + (size_t)n * m->kmax;
        for (int nz = m->ulevels_nod2D[n]; nz <= m->nlevels_nod2D[n] - 1; ++nz) {
            double vmax = aux[LAY(ne[0], nz) * 2 + 0];
            double vmin = aux[LAY(ne[0], nz) * 2 + 1];
            for (int k = 1; k < m->nod_in_elem2D_num[n]; ++k) {
                vmax = dmax(vmax, aux[LAY(ne[k], nz) * 2 + 0]);
                vmin = dmin(vmin, aux[LAY(ne[k], nz) * 2 + 1]);
            }
            if (g_mut == 49) {   /* the elements without level nz left out */
                vmax = -INFINITY;
                vmin = INFINITY;
                for (int k = 0; k < m->nod_in_elem2D_num[n]; ++k) {
                    const int el = ne[k];
                    if (nz < m->ulevels[el] || nz > m->nlevels[el] - 1) continue;
                    vmax = dmax(vmax, aux[LAY(el, nz) * 2 + 0]);
                    vmin = dmin(vmin, aux[LAY(el, nz) * 2 + 1]);
                }
            }
            tvmax[LAY(n, nz)] = vmax;
            tvmin[LAY(n, nz)] = vmin;
        }
    }
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nu1 = m->ulevels_nod2D[n], nl1 = m->nlevels_nod2D[n];
        fct_ttf_max[LAY(n, nu1)] = tvmax[LAY(n, nu1)] - lo[LAY(n, nu1)];
        fct_ttf_min[LAY(n, nu1)] = tvmin[LAY(n, nu1)] - lo[LAY(n, nu1)];
        for (int nz = nu1 + 1; nz <= nl1 - 2; ++nz) {
            fct_ttf_max[LAY(n, nz)] = dmax(dmax(tvmax[LAY(n, nz - 1)], tvmax[LAY(n, nz)]), tvmax[LAY(n, nz + 1)])
                                      - lo[LAY(n, nz)];
            fct_ttf_min[LAY(n, nz)] = dmin(dmin(tvmin[LAY(n, nz - 1)], tvmin[LAY(n, nz)]), tvmin[LAY(n, nz + 1)])
                                      - lo[LAY(n, nz)];
        }
        int nz = nl1 - 1;
        fct_ttf_max[LAY(n, nz)] = tvmax[LAY(n, nz)] - lo[LAY(n, nz)];
        fct_ttf_min[LAY(n, nz)] = tvmin[LAY(n, nz)] - lo[LAY(n, nz)];
    }
    /* b1 */
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n)
        for (int nz = m->ulevels_nod2D[n]; nz <= m->nlevels_nod2D[n] - 1; ++nz) {
            fct_plus[LAY(n, nz)] = 0.0;
            fct_minus[LAY(n, nz)] = 0.0;
        }
    if (g_mut != 18) {
        OMP_FOR
        for (int n = 0; n < m->nod2D; ++n)
            for (int nz = m->ulevels_nod2D[n]; nz <= m->nlevels_nod2D[n] - 1; ++nz) {
                fct_plus[LAY(n, nz)] = fct_plus[LAY(n, nz)]
                    + (dmax(0.0, adf_v[INT(n, nz)]) + dmax(0.0, -adf_v[INT(n, nz + 1)]));
                fct_minus[LAY(n, nz)] = fct_minus[LAY(n, nz)]
                    + (dmin(0.0, adf_v[INT(n, nz)]) + dmin(0.0, -adf_v[INT(n, nz + 1)]));
            }
    }
#ifdef _OPENMP
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n)
        for (int p = m->nod_edge_ptr[n]; p < m->nod_edge_ptr[n + 1]; ++p) {
            const int e = m->nod_edge_idx[p] >> 1, side = m->nod_edge_idx[p] & 1;
            edge_info q = get_edge(m, e);
            int nl12 = imax(q.nl1, q.nl2);
            int nu12 = q.nu1;
            if (q.nu2 > 0) nu12 = imin(q.nu1, q.nu2);
            for (int nz = nu12; nz <= nl12; ++nz) {
                const double f = side ? -adf_h[LAY(e, nz)] : adf_h[LAY(e, nz)];
                fct_plus[LAY(n, nz)] = fct_plus[LAY(n, nz)] + dmax(0.0, f);
                fct_minus[LAY(n, nz)] = fct_minus[LAY(n, nz)] + dmin(0.0, f);
            }
        }
#else
    for (int e = 0; e < m->edge2D; ++e) {
        edge_info q = get_edge(m, e);
        int nl12 = imax(q.nl1, q.nl2);
        int nu12 = q.nu1;
        if (q.nu2 > 0) nu12 = imin(q.nu1, q.nu2);
        for (int nz = nu12; nz <= nl12; ++nz) {
            fct_plus[LAY(q.n1, nz)] = fct_plus[LAY(q.n1, nz)] + dmax(0.0, adf_h[LAY(e, nz)]);
            fct_minus[LAY(q.n1, nz)] = fct_minus[LAY(q.n1, nz)] + dmin(0.0, adf_h[LAY(e, nz)]);
        }
        for (int nz = nu12; nz <= nl12; ++nz) {
            fct_plus[LAY(q.n2, nz)] = fct_plus[LAY(q.n2, nz)] + dmax(0.0, -adf_h[LAY(e, nz)]);
            fct_minus[LAY(q.n2, nz)] = fct_minus[LAY(q.n2, nz)] + dmin(0.0, -adf_h[LAY(e, nz)]);
        }
    }
#endif
    /* b2 */
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n)
        for (int nz = m->ulevels_nod2D[n]; nz <= m->nlevels_nod2D[n] - 1; ++nz) {
            double flux = fct_plus[LAY(n, nz)] * dt / m->areasvol[INT(n, nz)] / m->hnode_new[LAY(n, nz)] + flux_eps;
            fct_plus[LAY(n, nz)] = dmin(1.0, fct_ttf_max[LAY(n, nz)] / flux);
            flux = fct_minus[LAY(n, nz)] * dt / m->areasvol[INT(n, nz)] / m->hnode_new[LAY(n, nz)] - flux_eps;
            fct_minus[LAY(n, nz)] = dmin(1.0, fct_ttf_min[LAY(n, nz)] / flux);
        }
    /* b3 vertical */
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n) {
        int nu1 = m->ulevels_nod2D[n], nl1 = m->nlevels_nod2D[n];
        int nz = nu1;
        double ae = 1.0;
        double flux = adf_v[INT(n, nz)];
        if (flux >= 0.0) ae = dmin(ae, fct_plus[LAY(n, nz)]);
        else ae = dmin(ae, fct_minus[LAY(n, nz)]);
        adf_v[INT(n, nz)] = ae * adf_v[INT(n, nz)];
        for (nz = nu1 + 1; nz <= nl1 - 1; ++nz) {
            ae = 1.0;
            flux = adf_v[INT(n, nz)];
            if (flux >= 0.0) {
                ae = dmin(ae, (g_mut == 8) ? fct_plus[LAY(n, nz - 1)] : fct_minus[LAY(n, nz - 1)]);
                ae = dmin(ae, fct_plus[LAY(n, nz)]);
            } else {
                ae = dmin(ae, fct_plus[LAY(n, nz - 1)]);
                ae = dmin(ae, fct_minus[LAY(n, nz)]);
            }
            adf_v[INT(n, nz)] = ae * adf_v[INT(n, nz)];
        }
    }
    /* b3 horizontal */
    OMP_FOR
    for (int e = 0; e < m->edge2D; ++e) {
        edge_info q = get_edge(m, e);
        int nl12 = imax(q.nl1, q.nl2);
        int nu12 = q.nu1;
        if (q.nu2 > 0) nu12 = imin(q.nu1, q.nu2);
        for (int nz = nu12; nz <= nl12; ++nz) {
            double ae = 1.0;
            double flux = adf_h[LAY(e, nz)];
            if (flux >= 0.0) {
                ae = dmin(ae, fct_plus[LAY(q.n1, nz)]);
                ae = dmin(ae, fct_minus[LAY(q.n2, nz)]);
            } else {
                ae = dmin(ae, fct_minus[LAY(q.n1, nz)]);
                ae = dmin(ae, fct_plus[LAY(q.n2, nz)]);
            }
            adf_h[LAY(e, nz)] = ae * adf_h[LAY(e, nz)];
        }
    }
    free(aux);
    free(tvmax);
    free(tvmin);
}

/* ------------------------------------------------------------------------ */
/* oce_tra_adv_flux2dtracer: oce_adv_tra_driver.F90:494-646                  */
/* ------------------------------------------------------------------------ */
void fo_oce_tra_adv_flux2dtracer(const fo_mesh *m, double dt, double *dttf_h, double *dttf_v,
                                 const double *flux_h, const double *flux_v,
                                 int use_lo, const double *ttf, const double *lo)
{
    const int L = m->nl - 1, NL = m->nl;
    if (use_lo) {
        OMP_FOR
        for (int n = 0; n < m->nod2D; ++n)
            for (int nz = m->ulevels_nod2D[n]; nz <= m->nlevels_nod2D[n] - 1; ++nz)
                dttf_v[LAY(n, nz)] = dttf_v[LAY(n, nz)] - ttf[LAY(n, nz)] * m->hnode[LAY(n, nz)]
                                     + lo[LAY(n, nz)] * m->hnode_new[LAY(n, nz)];
    }
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n)
        for (int nz = m->ulevels_nod2D[n]; nz <= m->nlevels_nod2D[n] - 1; ++nz)
            dttf_v[LAY(n, nz)] = dttf_v[LAY(n, nz)]
                + (flux_v[INT(n, nz)] - flux_v[INT(n, nz + 1)]) * dt / m->areasvol[INT(n, nz)];
#ifdef _OPENMP
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n)
        for (int p = m->nod_edge_ptr[n]; p < m->nod_edge_ptr[n + 1]; ++p) {
            const int e = m->nod_edge_idx[p] >> 1, side = m->nod_edge_idx[p] & 1;
            edge_info q = get_edge(m, e);
            int nl12 = imax(q.nl1, q.nl2);
            int nu12 = q.nu1;
            if (q.nu2 > 0) nu12 = imin(q.nu1, q.nu2);
            for (int nz = nu12; nz <= nl12; ++nz) {
                if (side) dttf_h[LAY(n, nz)] = dttf_h[LAY(n, nz)] - flux_h[LAY(e, nz)] * dt / m->areasvol[INT(n, nz)];
                else dttf_h[LAY(n, nz)] = dttf_h[LAY(n, nz)] + flux_h[LAY(e, nz)] * dt / m->areasvol[INT(n, nz)];
            }
        }
#else
    for (int e = 0; e < m->edge2D; ++e) {
        edge_info q = get_edge(m, e);
        int nl12 = imax(q.nl1, q.nl2);
        int nu12 = q.nu1;
        if (q.nu2 > 0) nu12 = imin(q.nu1, q.nu2);
        for (int nz = nu12; nz <= nl12; ++nz)
            dttf_h[LAY(q.n1, nz)] = dttf_h[LAY(q.n1, nz)] + flux_h[LAY(e, nz)] * dt / m->areasvol[INT(q.n1, nz)];
        int nb = (g_mut == 10) ? q.n1 : q.n2;
        for (int nz = nu12; nz <= nl12; ++nz)
            dttf_h[LAY(q.n2, nz)] = dttf_h[LAY(q.n2, nz)] - flux_h[LAY(e, nz)] * dt / m->areasvol[INT(nb, nz)];
    }
#endif
}

/* ------------------------------------------------------------------------ */
/* adv_tra_ver_cdiff: oce_adv_tra_ver.F90:635-695                            */
/* ------------------------------------------------------------------------ */
void fo_adv_tra_ver_cdiff(const fo_mesh *m, const double *W, const double *ttf, double *flux, int init_zero)
{
    const int L = m->nl - 1, NL = m->nl;
    double *tvert = (double *)malloc(sizeof(double) * (size_t)(NL + 1));
    if (init_zero) memset(flux, 0, (size_t)m->nod2D * NL * sizeof(double));
    for (int n = 0; n < m->nod2D; ++n) {
        const int nzmax = m->nlevels_nod2D[n] - 1, nzmin = m->ulevels_nod2D[n];
        tvert[nzmin] = -W[INT(n, nzmin)] * ttf[LAY(n, nzmin)] * m->area[INT(n, nzmin)];
        tvert[nzmax + 1] = 0.0;
        for (int nz = nzmin + 1; nz <= nzmax; ++nz) {
            const double tv = (g_mut == 25) ? ttf[LAY(n, nz)] : 0.5 * (ttf[LAY(n, nz - 1)] + ttf[LAY(n, nz)]);
            tvert[nz] = -tv * W[INT(n, nz)] * m->area[INT(n, nz)];
        }
        for (int nz = nzmin; nz <= nzmax; ++nz) flux[INT(n, nz)] = tvert[nz] - flux[INT(n, nz)];
    }
    free(tvert);
    (void)L;
}

/* ------------------------------------------------------------------------ */
/* adv_tra_vert_ppm: oce_adv_tra_ver.F90:438-631 (Colella and Woodward 1984: */
/* interface values from eqs. 1.6-1.8 on the new thicknesses, limited        */
/* parabolas, swept-volume fluxes with the old thicknesses)                  */
/* ------------------------------------------------------------------------ */
static double fo_sign1(double b) { return signbit(b) ? -1.0 : 1.0; }

void fo_adv_tra_vert_ppm(const fo_mesh *m, double dt, const double *W, const double *ttf, double *flux,
                         int init_zero)
{
    const int L = m->nl - 1, NL = m->nl;
    double *tvert = (double *)malloc(sizeof(double) * (size_t)(NL + 1));
    double *tv = (double *)malloc(sizeof(double) * (size_t)(NL + 1));
    if (init_zero) memset(flux, 0, (size_t)m->nod2D * NL * sizeof(double));
    for (int n = 0; n < m->nod2D; ++n) {
        const int nzmax = m->nlevels_nod2D[n], nzmin = m->ulevels_nod2D[n];
#define TT(k) ttf[LAY(n, (k))]
#define HN(k) m->hnode_new[LAY(n, (k))]
        tv[nzmin] = TT(nzmin);
        tv[nzmin + 1] = 0.5 * (TT(nzmin) + TT(nzmin + 1));
        tv[nzmax - 1] = 0.5 * (TT(nzmax - 2) + TT(nzmax - 1));
        tv[nzmax] = TT(nzmax - 1);
        for (int nz = nzmin + 1; nz <= nzmax - 3; ++nz) {
            const double dzjm1 = HN(nz - 1), dzj = HN(nz), dzjp1 = HN(nz + 1), dzjp2 = HN(nz + 2);
            double deltaj = dzj / (dzjm1 + dzj + dzjp1) *
                            ((2.0 * dzjm1 + dzj) / (dzjp1 + dzj) * (TT(nz + 1) - TT(nz)) +
                             (dzj + 2.0 * dzjp1) / (dzjm1 + dzj) * (TT(nz) - TT(nz - 1)));
            double deltajp1 = dzjp1 / (dzj + dzjp1 + dzjp2) *
                              ((2.0 * dzj + dzjp1) / (dzjp2 + dzjp1) * (TT(nz + 2) - TT(nz + 1)) +
                               (dzjp1 + 2.0 * dzjp2) / (dzj + dzjp1) * (TT(nz + 1) - TT(nz)));
            if ((TT(nz + 1) - TT(nz)) * (TT(nz) - TT(nz - 1)) > 0.0) {
                double a = fabs(deltaj), b = 2.0 * fabs(TT(nz + 1) - TT(nz)), c = 2.0 * fabs(TT(nz) - TT(nz - 1));
                deltaj = dmin(dmin(a, b), c) * fo_sign1(deltaj);
            } else {
                deltaj = 0.0;
            }
            if ((TT(nz + 2) - TT(nz + 1)) * (TT(nz + 1) - TT(nz)) > 0.0) {
                double a = fabs(deltajp1), b = 2.0 * fabs(TT(nz + 2) - TT(nz + 1)), c = 2.0 * fabs(TT(nz + 1) - TT(nz));
                deltajp1 = dmin(dmin(a, b), c) * fo_sign1(deltajp1);
            } else {
                deltajp1 = 0.0;
            }
            tv[nz + 1] = TT(nz) + dzj / (dzj + dzjp1) * (TT(nz + 1) - TT(nz)) +
                         1.0 / (dzjm1 + dzj + dzjp1 + dzjp2) *
                             (2.0 * dzjp1 * dzj / (dzj + dzjp1) *
                                  ((dzjm1 + dzj) / (2.0 * dzj + dzjp1) - (dzjp2 + dzjp1) / (2.0 * dzjp1 + dzj)) *
                                  (TT(nz + 1) - TT(nz)) -
                              dzj * (dzjm1 + dzj) / (2.0 * dzj + dzjp1) * deltajp1 +
                              dzjp1 * (dzjp1 + dzjp2) / (dzj + 2.0 * dzjp1) * deltaj);
        }
        for (int k = 1; k <= nzmax; ++k) tvert[k] = 0.0;
        for (int nz = nzmin; nz <= nzmax - 1; ++nz) {
            const double wt = W[INT(n, nz)], wb = W[INT(n, nz + 1)];
            if ((wt <= 0.0) && (wb >= 0.0)) continue;
            double aL = tv[nz], aR = tv[nz + 1];
            const double t = TT(nz);
            if ((aR - t) * (t - aL) <= 0.0) { aL = t; aR = t; }
            if ((aR - aL) * (t - 0.5 * (aL + aR)) > (aR - aL) * (aR - aL) / 6.0) aL = 3.0 * t - 2.0 * aR;
            if ((aR - aL) * (t - 0.5 * (aR + aL)) < -((aR - aL) * (aR - aL) / 6.0)) aR = 3.0 * t - 2.0 * aL;
            const double dzj = m->hnode[LAY(n, nz)];
            const double aj = 6.0 * (t - 0.5 * (aL + aR));
            if (wt > 0.0) {
                const double x = dmin(wt * dt / dzj, 1.0);
                tvert[nz] = (g_mut == 24) ? (-aR - 0.5 * x * (aL - aR + (1.0 - 2.0 / 3.0 * x) * aj))
                                          : (-aL - 0.5 * x * (aR - aL + (1.0 - 2.0 / 3.0 * x) * aj));
                tvert[nz] = tvert[nz] * m->area[INT(n, nz)] * wt;
            }
            if (wb < 0.0) {
                const double x = dmin(-wb * dt / dzj, 1.0);
                tvert[nz + 1] = (-aR + 0.5 * x * (aR - aL - (1.0 - 2.0 / 3.0 * x) * aj));
                tvert[nz + 1] = tvert[nz + 1] * m->area[INT(n, nz + 1)] * wb;
            }
        }
        tvert[nzmin] = -tv[nzmin] * W[INT(n, nzmin)] * m->area[INT(n, nzmin)];
        tvert[nzmax] = 0.0;
        for (int nz = nzmin; nz <= nzmax; ++nz) flux[INT(n, nz)] = tvert[nz] - flux[INT(n, nz)];
#undef TT
#undef HN
    }
    free(tv);
    free(tvert);
    (void)L;
}

/* ------------------------------------------------------------------------ */
/* adv_tra_vert_impl: oce_adv_tra_ver.F90:90-240. Implicit first-order      */
/* upwind vertical advection of ttf (in place) with the implicit part w_i   */
/* of the split vertical velocity; one tridiagonal solve per column.         */
/* ------------------------------------------------------------------------ */
void fo_adv_tra_vert_impl(const fo_mesh *m, double dt, const double *W, double *ttf)
{
    const int L = m->nl - 1, NL = m->nl;
    double *a = (double *)calloc(NL, sizeof(double)), *b = (double *)calloc(NL, sizeof(double));
    double *c = (double *)calloc(NL, sizeof(double)), *tr = (double *)calloc(NL, sizeof(double));
    double *cp = (double *)calloc(NL, sizeof(double)), *tp = (double *)calloc(NL, sizeof(double));
    for (int n = 0; n < m->nod2D; ++n) {
        const int nzmax = m->nlevels_nod2D[n], nzmin = m->ulevels_nod2D[n];
#define WI(k) W[INT(n, (k))]
#define AR(k) m->area[INT(n, (k))]
#define AV(k) m->areasvol[INT(n, (k))]
#define HN(k) m->hnode_new[LAY(n, (k))]
#define TF(k) ttf[LAY(n, (k))]
#define I(k) ((k) - 1)
        for (int i = 0; i < NL; ++i) { a[i] = b[i] = c[i] = tr[i] = tp[i] = cp[i] = 0.0; }
        const double zinv = 1.0 * dt;
        int nz = nzmin;
        double v_adv = zinv * AR(nz) / AV(nz);
        a[I(nz)] = 0.0;
        b[I(nz)] = HN(nz) + WI(nz) * v_adv;
        v_adv = zinv * AR(nz + 1) / AV(nz);
        b[I(nz)] = b[I(nz)] - dmin(0.0, WI(nz + 1)) * v_adv;
        c[I(nz)] = -dmax(0.0, WI(nz + 1)) * v_adv;
        for (nz = nzmin + 1; nz <= nzmax - 2; ++nz) {
            v_adv = zinv * AR(nz) / AV(nz);
            a[I(nz)] = (g_mut == 27 ? dmax(0.0, WI(nz)) : dmin(0.0, WI(nz))) * v_adv;
            b[I(nz)] = HN(nz) + dmax(0.0, WI(nz)) * v_adv;
            v_adv = zinv * AR(nz + 1) / AV(nz);
            b[I(nz)] = b[I(nz)] - dmin(0.0, WI(nz + 1)) * v_adv;
            c[I(nz)] = -dmax(0.0, WI(nz + 1)) * v_adv;
        }
        nz = nzmax - 1;
        v_adv = zinv * AR(nz) / AV(nz);
        a[I(nz)] = dmin(0.0, WI(nz)) * v_adv;
        b[I(nz)] = HN(nz) + dmax(0.0, WI(nz)) * v_adv;
        c[I(nz)] = 0.0;
        nz = nzmin;
        double dz = HN(nz);
        tr[I(nz)] = -(b[I(nz)] - dz) * TF(nz) - c[I(nz)] * TF(nz + 1);
        for (nz = nzmin + 1; nz <= nzmax - 2; ++nz) {
            dz = HN(nz);
            tr[I(nz)] = -a[I(nz)] * TF(nz - 1) - (b[I(nz)] - dz) * TF(nz) - c[I(nz)] * TF(nz + 1);
        }
        nz = nzmax - 1;
        dz = HN(nz);
        tr[I(nz)] = -a[I(nz)] * TF(nz - 1) - (b[I(nz)] - dz) * TF(nz);
        cp[I(nzmin)] = c[I(nzmin)] / b[I(nzmin)];
        tp[I(nzmin)] = tr[I(nzmin)] / b[I(nzmin)];
        for (nz = nzmin + 1; nz <= nzmax - 1; ++nz) {
            const double mm = b[I(nz)] - cp[I(nz - 1)] * a[I(nz)];
            cp[I(nz)] = c[I(nz)] / mm;
            tp[I(nz)] = (tr[I(nz)] - tp[I(nz - 1)] * a[I(nz)]) / mm;
        }
        tr[I(nzmax - 1)] = tp[I(nzmax - 1)];
        for (nz = nzmax - 2; nz >= nzmin; --nz) tr[I(nz)] = tp[I(nz)] - cp[I(nz)] * tr[I(nz + 1)];
        for (nz = nzmin; nz <= nzmax - 1; ++nz) TF(nz) = TF(nz) + tr[I(nz)];
#undef WI
#undef AR
#undef AV
#undef HN
#undef TF
#undef I
    }
    free(a); free(b); free(c); free(tr); free(cp); free(tp);
    (void)L;
}

/* ------------------------------------------------------------------------ */
/* do_oce_adv_tra: oce_adv_tra_driver.F90:46-490 (ltra_diag and DVD off)     */
/* ------------------------------------------------------------------------ */
int fo_do_oce_adv_tra(const fo_mesh *m, const fo_dyn *d, double dt, fo_tracer *t, fo_work *w)
{
    const int L = m->nl - 1, NL = m->nl;
    if (t->hor != FO_HOR_MFCT && t->hor != FO_HOR_UPW1 && t->hor != FO_HOR_MUSCL) return -3;
    if (t->ver != FO_VER_QR4C && t->ver != FO_VER_UPW1 && t->ver != FO_VER_CDIFF && t->ver != FO_VER_PPM) return -4;
    int fct = (t->lim == FO_LIM_FCT);
    if (fct) {
        fo_adv_tra_hor_upw1(m, d->uv, t->values, w->adv_flux_hor, 1);
        memset(w->fct_LO, 0, (size_t)m->nod2D * L * sizeof(double));
#ifdef _OPENMP
        OMP_FOR
        for (int n = 0; n < m->nod2D; ++n)
            for (int p = m->nod_edge_ptr[n]; p < m->nod_edge_ptr[n + 1]; ++p) {
                const int e = m->nod_edge_idx[p] >> 1, side = m->nod_edge_idx[p] & 1;
                edge_info q = get_edge(m, e);
                int nl12 = imax(q.nl1, q.nl2);
                int nu12 = q.nu1;
                if (q.nu2 > 0) nu12 = imin(q.nu1, q.nu2);
                for (int nz = nu12; nz <= nl12; ++nz) {
                    if (side) w->fct_LO[LAY(n, nz)] = w->fct_LO[LAY(n, nz)] - w->adv_flux_hor[LAY(e, nz)];
                    else w->fct_LO[LAY(n, nz)] = w->fct_LO[LAY(n, nz)] + w->adv_flux_hor[LAY(e, nz)];
                }
            }
#else
        for (int e = 0; e < m->edge2D; ++e) {
            edge_info q = get_edge(m, e);
            int nl12 = imax(q.nl1, q.nl2);
            int nu12 = q.nu1;
            if (q.nu2 > 0) nu12 = imin(q.nu1, q.nu2);
            for (int nz = nu12; nz <= nl12; ++nz) {
                w->fct_LO[LAY(q.n1, nz)] = w->fct_LO[LAY(q.n1, nz)] + w->adv_flux_hor[LAY(e, nz)];
                w->fct_LO[LAY(q.n2, nz)] = w->fct_LO[LAY(q.n2, nz)] - w->adv_flux_hor[LAY(e, nz)];
            }
        }
#endif
        fo_adv_tra_ver_upw1(m, d->w_e, t->values, w->adv_flux_ver, 1);
        OMP_FOR
        for (int n = 0; n < m->nod2D; ++n)
            for (int nz = m->ulevels_nod2D[n]; nz <= m->nlevels_nod2D[n] - 1; ++nz)
                w->fct_LO[LAY(n, nz)] = (t->values[LAY(n, nz)] * m->hnode[LAY(n, nz)]
                    + (w->fct_LO[LAY(n, nz)] + (w->adv_flux_ver[INT(n, nz)] - w->adv_flux_ver[INT(n, nz + 1)]))
                      * dt / m->areasvol[INT(n, nz)]) / m->hnode_new[LAY(n, nz)];
        if (d->use_wsplit) {   /* oce_adv_tra_driver.F90:323-331 */
            fo_adv_tra_vert_impl(m, dt, d->w_i, w->fct_LO);
            fo_adv_tra_ver_upw1(m, d->w, t->values, w->adv_flux_ver, 1);
        }
    }
    int zero = !fct;
    if (t->hor == FO_HOR_MFCT) {
        fo_adv_tra_hor_mfct(m, d->uv, t->valuesAB, t->ph, w->adv_flux_hor, w->edge_up_dn_grad, zero);
    } else if (t->hor == FO_HOR_MUSCL) {
        int *nblay = (int *)malloc(sizeof(int) * (size_t)m->nod2D);
        if (!nblay) return -5;
        fo_muscl_nboundary_lay(m, nblay);
        fo_adv_tra_hor_muscl(m, d->uv, t->valuesAB, t->ph, w->adv_flux_hor, w->edge_up_dn_grad, nblay, zero);
        free(nblay);
    }
    else
        fo_adv_tra_hor_upw1(m, d->uv, t->valuesAB, w->adv_flux_hor, zero);
    const double *pw = fct ? d->w : d->w_e;
    if (t->ver == FO_VER_QR4C)
        fo_adv_tra_ver_qr4c(m, pw, t->valuesAB, t->pv, w->adv_flux_ver, zero);
    else if (t->ver == FO_VER_CDIFF)
        fo_adv_tra_ver_cdiff(m, pw, t->valuesAB, w->adv_flux_ver, zero);
    else if (t->ver == FO_VER_PPM)
        fo_adv_tra_vert_ppm(m, dt, pw, t->valuesAB, w->adv_flux_ver, zero);
    else
        fo_adv_tra_ver_upw1(m, pw, t->valuesAB, w->adv_flux_ver, zero);
    if (fct) {
        fo_oce_tra_adv_fct(m, dt, t->values, w->fct_LO, w->adv_flux_hor, w->adv_flux_ver,
                           w->fct_ttf_min, w->fct_ttf_max, w->fct_plus, w->fct_minus);
        fo_oce_tra_adv_flux2dtracer(m, dt, w->del_ttf_advhoriz, w->del_ttf_advvert,
                                    w->adv_flux_hor, w->adv_flux_ver, 1, t->values, w->fct_LO);
    } else {
        fo_oce_tra_adv_flux2dtracer(m, dt, w->del_ttf_advhoriz, w->del_ttf_advvert,
                                    w->adv_flux_hor, w->adv_flux_ver, 0, NULL, NULL);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* T* update: oce_ale_tracer.F90:318-322 and 742-755                         */
/* ------------------------------------------------------------------------ */
void fo_diff_tracers_ale_tstar(const fo_mesh *m, const fo_dyn *d, double dt, fo_tracer *t, fo_work *w)
{
    const int L = m->nl - 1;
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n)
        for (int nz = 1; nz <= L; ++nz)
            w->del_ttf[LAY(n, nz)] = w->del_ttf[LAY(n, nz)] + w->del_ttf_advhoriz[LAY(n, nz)]
                                     + w->del_ttf_advvert[LAY(n, nz)];
    /* diff_tracers_ale (oce_ale_tracer.F90:595-626): horizontal (and Redi)
       diffusion, explicit vertical diffusion when i_vert_diff is off, the
       explicit vertical Redi part; Ki = 0 makes the first a no-op */
    if (d && d->Ki) fo_diff_part_hor_redi(m, d, dt, w->tr_xy, w->tr_z, w->del_ttf);
    if (d && !t->i_vert_diff) fo_diff_ver_part_expl_ale(m, d, dt, t, w->del_ttf);
    if (d && d->redi && d->Ki && d->slope_tapered) fo_diff_ver_part_redi_expl(m, d, dt, w->tr_xy, w->del_ttf);
    OMP_FOR
    for (int n = 0; n < m->nod2D; ++n)
        for (int nz = m->ulevels_nod2D[n]; nz <= m->nlevels_nod2D[n] - 1; ++nz) {
            w->del_ttf[LAY(n, nz)] = w->del_ttf[LAY(n, nz)]
                + t->values[LAY(n, nz)] * (m->hnode[LAY(n, nz)] - m->hnode_new[LAY(n, nz)]);
            t->values[LAY(n, nz)] = t->values[LAY(n, nz)] + w->del_ttf[LAY(n, nz)] / m->hnode_new[LAY(n, nz)];
        }
}

/* ------------------------------------------------------------------------ */
/* diff_part_hor_redi: oce_ale_tracer.F90:1526-1689. Per edge and level the */
/* flux K (grad T + isredi S T_z) through the edge's dual segments, cases    */
/* (A)-(E) for levels where one or both elements exist; scattered to both   */
/* nodes in edge order.                                                      */
/* ------------------------------------------------------------------------ */
void fo_diff_part_hor_redi(const fo_mesh *m, const fo_dyn *d, double dt, const double *tr_xy, const double *tr_z,
                           double *del_ttf)
{
    const int L = m->nl - 1, NL = m->nl;
    const double isredi = d->redi ? 1.0 : 0.0;
    double *rhs1 = (double *)calloc(L + 1, sizeof(double)), *rhs2 = (double *)calloc(L + 1, sizeof(double));
    for (int edge = 0; edge < m->edge2D; ++edge) {
        for (int i = 0; i <= L; ++i) rhs1[i] = rhs2[i] = 0.0;
        const double dX1 = m->edge_cross_dxdy[4 * edge + 0], dY1 = m->edge_cross_dxdy[4 * edge + 1];
        const int el1 = m->edge_tri[2 * edge], el2 = m->edge_tri[2 * edge + 1];
        const int n1 = m->edges[2 * edge], n2 = m->edges[2 * edge + 1];
        const int nl1 = m->nlevels[el1] - 1, ul1 = m->ulevels[el1];
        int nl2 = 0, ul2 = 0;
        double dX2 = 0.0, dY2 = 0.0;
        if (el2 >= 0) {
            nl2 = m->nlevels[el2] - 1;
            ul2 = m->ulevels[el2];
            dX2 = m->edge_cross_dxdy[4 * edge + 2];
            dY2 = m->edge_cross_dxdy[4 * edge + 3];
        }
        const int nl12 = imin(nl1, nl2), ul12 = imax(ul1, ul2);
#define TXY(c, nz, el) tr_xy[((size_t)(el) * L + (nz) - 1) * 2 + (c)]
        /* Kh, Fx, Fy at level nz from the gradient (Tx, Ty) */
#define FLUX(nz)                                                                                            \
    const double Kh = (g_mut == 31 ? 2.0 * d->Ki[LAY(n1, nz)] : d->Ki[LAY(n1, nz)] + d->Ki[LAY(n2, nz)]) / 2.0; \
    double SxTz = 0.0, SyTz = 0.0;                                                                          \
    if (d->slope_tapered) {                                                                                 \
        const double Tz1 = 0.5 * (tr_z[INT(n1, nz)] + tr_z[INT(n1, (nz) + 1)]);                             \
        const double Tz2 = 0.5 * (tr_z[INT(n2, nz)] + tr_z[INT(n2, (nz) + 1)]);                             \
        SxTz = (Tz1 * d->slope_tapered[LAY(n1, nz) * 3 + 0] + Tz2 * d->slope_tapered[LAY(n2, nz) * 3 + 0]) / 2.0; \
        SyTz = (Tz1 * d->slope_tapered[LAY(n1, nz) * 3 + 1] + Tz2 * d->slope_tapered[LAY(n2, nz) * 3 + 1]) / 2.0; \
    }                                                                                                       \
    const double Fx = Kh * (Tx + SxTz * isredi);                                                            \
    const double Fy = Kh * (Ty + SyTz * isredi);
        for (int nz = ul1; nz <= ul12 - 1; ++nz) {   /* (A) */
            const double dz = m->helem[LAY(el1, nz)];
            const double Tx = TXY(0, nz, el1), Ty = TXY(1, nz, el1);
            FLUX(nz)
            const double c = (-dX1 * Fy + dY1 * Fx) * dz;
            rhs1[nz] = rhs1[nz] + c;
            rhs2[nz] = rhs2[nz] - c;
        }
        if (ul2 > 0)
            for (int nz = ul2; nz <= ul12 - 1; ++nz) {   /* (B) */
                const double dz = m->helem[LAY(el2, nz)];
                const double Tx = TXY(0, nz, el2), Ty = TXY(1, nz, el2);
                FLUX(nz)
                const double c = (dX2 * Fy - dY2 * Fx) * dz;
                rhs1[nz] = rhs1[nz] + c;
                rhs2[nz] = rhs2[nz] - c;
            }
        for (int nz = ul12; nz <= nl12; ++nz) {   /* (C) */
            const double dz = (m->helem[LAY(el1, nz)] + m->helem[LAY(el2, nz)]) / 2.0;
            const double Tx = 0.5 * (TXY(0, nz, el1) + TXY(0, nz, el2));
            const double Ty = 0.5 * (TXY(1, nz, el1) + TXY(1, nz, el2));
            FLUX(nz)
            const double dX2m = (g_mut == 30) ? -dX2 : dX2, dY2m = (g_mut == 30) ? -dY2 : dY2;
            const double c = ((dX2m - dX1) * Fy - (dY2m - dY1) * Fx) * dz;
            rhs1[nz] = rhs1[nz] + c;
            rhs2[nz] = rhs2[nz] - c;
        }
        for (int nz = nl12 + 1; nz <= nl1; ++nz) {   /* (D) */
            const double dz = m->helem[LAY(el1, nz)];
            const double Tx = TXY(0, nz, el1), Ty = TXY(1, nz, el1);
            FLUX(nz)
            const double c = (-dX1 * Fy + dY1 * Fx) * dz;
            rhs1[nz] = rhs1[nz] + c;
            rhs2[nz] = rhs2[nz] - c;
        }
        for (int nz = nl12 + 1; nz <= nl2; ++nz) {   /* (E) */
            const double dz = m->helem[LAY(el2, nz)];
            const double Tx = TXY(0, nz, el2), Ty = TXY(1, nz, el2);
            FLUX(nz)
            const double c = (dX2 * Fy - dY2 * Fx) * dz;
            rhs1[nz] = rhs1[nz] + c;
            rhs2[nz] = rhs2[nz] - c;
        }
#undef FLUX
#undef TXY
        const int nls = imax(nl1, nl2);
        int uls = ul1;
        if (ul2 > 0) uls = imin(ul1, ul2);
        for (int nz = uls; nz <= nls; ++nz)
            del_ttf[LAY(n1, nz)] = del_ttf[LAY(n1, nz)] + rhs1[nz] * dt / m->areasvol[INT(n1, nz)];
        for (int nz = uls; nz <= nls; ++nz)
            del_ttf[LAY(n2, nz)] = del_ttf[LAY(n2, nz)] + rhs2[nz] * dt / m->areasvol[INT(n2, nz)];
    }
    free(rhs1);
    free(rhs2);
    (void)NL;
}

/* ------------------------------------------------------------------------ */
/* diff_ver_part_expl_ale: oce_ale_tracer.F90:841-902 (i_vert_diff off)      */
/* ------------------------------------------------------------------------ */
void fo_diff_ver_part_expl_ale(const fo_mesh *m, const fo_dyn *d, double dt, const fo_tracer *t, double *del_ttf)
{
    const int L = m->nl - 1, NL = m->nl;
    double *vd = (double *)calloc(NL + 1, sizeof(double));
    for (int n = 0; n < m->nod2D; ++n) {
        const int nl1 = m->nlevels_nod2D[n] - 1, ul1 = m->ulevels_nod2D[n];
        for (int i = 0; i <= NL; ++i) vd[i] = 0.0;
        double flux;
        if (t->ID == 1) flux = -d->heat_flux[n] / VCPW;
        else if (t->ID == 2) flux = d->virtual_salt[n] + d->relax_salt[n] + d->real_salt_flux[n] * d->is_nonlinfs;
        else flux = 0.0;
        vd[ul1] = flux;
        for (int nz = ul1 + 1; nz <= nl1; ++nz) {
            const double zinv1 = (g_mut == 29) ? 1.0 / (m->zbar_3d_n[INT(n, nz - 1)] - m->zbar_3d_n[INT(n, nz)])
                                               : 1.0 / (m->Z_3d_n[LAY(n, nz - 1)] - m->Z_3d_n[LAY(n, nz)]);
            vd[nz] = d->Kv[INT(n, nz)] * (t->values[LAY(n, nz - 1)] - t->values[LAY(n, nz)]) * zinv1
                     * m->area[INT(n, nz)];
        }
        for (int nz = ul1; nz <= nl1 - 1; ++nz)
            del_ttf[LAY(n, nz)] = del_ttf[LAY(n, nz)] + (vd[nz] - vd[nz + 1])
                / (m->zbar_3d_n[INT(n, nz)] - m->zbar_3d_n[INT(n, nz + 1)]) * dt / m->areasvol[INT(n, nz)];
        del_ttf[LAY(n, nl1)] = del_ttf[LAY(n, nl1)]
            + (vd[nl1] / (m->zbar_3d_n[INT(n, nl1)] - m->zbar_3d_n[INT(n, nl1 + 1)])) * dt / m->areasvol[INT(n, nl1)];
    }
    free(vd);
    (void)L;
}

/* ------------------------------------------------------------------------ */
/* diff_ver_part_redi_expl: oce_ale_tracer.F90:1439-1522. Node gradients    */
/* (element gradients weighted by area / 3 / areasvol), then the vertical    */
/* flux Ki (S . grad T) on the old thicknesses.                              */
/* ------------------------------------------------------------------------ */
void fo_diff_ver_part_redi_expl(const fo_mesh *m, const fo_dyn *d, double dt, const double *tr_xy,
                                double *del_ttf)
{
    const int L = m->nl - 1, NL = m->nl;
    double *txyn = (double *)calloc((size_t)m->nod2D * L * 2, sizeof(double));
    for (int n = 0; n < m->nod2D; ++n) {
        const int nl1 = m->nlevels_nod2D[n] - 1, ul1 = m->ulevels_nod2D[n];
        for (int nz = ul1; nz <= nl1; ++nz) {
            double Tx = 0.0, Ty = 0.0;
            for (int k = 0; k < m->nod_in_elem2D_num[n]; ++k) {
                const int elem = m->nod_in_elem2D[(size_t)n * m->kmax + k];
                if (nz <= m->nlevels[elem] - 1 && nz >= m->ulevels[elem]) {
                    Tx = Tx + tr_xy[((size_t)elem * L + nz - 1) * 2 + 0] * m->elem_area[elem];
                    Ty = Ty + tr_xy[((size_t)elem * L + nz - 1) * 2 + 1] * m->elem_area[elem];
                }
            }
            txyn[LAY(n, nz) * 2 + 0] = Tx / 3.0 / m->areasvol[INT(n, nz)];
            txyn[LAY(n, nz) * 2 + 1] = Ty / 3.0 / m->areasvol[INT(n, nz)];
        }
    }
    double *vd = (double *)calloc(NL + 2, sizeof(double)), *zbar_n = (double *)calloc(NL + 2, sizeof(double));
    double *z_n = (double *)calloc(NL + 2, sizeof(double));
    const double *S = d->slope_tapered;
#define SG(nz) (S[LAY(n, nz) * 3 + 0] * txyn[LAY(n, nz) * 2 + 0] + S[LAY(n, nz) * 3 + 1] * txyn[LAY(n, nz) * 2 + 1])
    for (int n = 0; n < m->nod2D; ++n) {
        const int nl1 = m->nlevels_nod2D[n] - 1, ul1 = m->ulevels_nod2D[n];
        for (int i = 0; i <= NL + 1; ++i) vd[i] = zbar_n[i] = z_n[i] = 0.0;
        zbar_n[nl1 + 1] = m->zbar_n_bot[n];
        z_n[nl1] = zbar_n[nl1 + 1] + m->hnode[LAY(n, nl1)] / 2.0;
        for (int nz = nl1; nz >= ul1 + 1; --nz) {
            zbar_n[nz] = zbar_n[nz + 1] + m->hnode[LAY(n, nz)];
            z_n[nz - 1] = zbar_n[nz] + m->hnode[LAY(n, nz - 1)] / 2.0;
        }
        zbar_n[ul1] = zbar_n[ul1 + 1] + m->hnode[LAY(n, ul1)];
        for (int nz = ul1 + 1; nz <= nl1; ++nz) {
            const double wup = (g_mut == 32) ? (z_n[nz - 1] - zbar_n[nz + 1]) : (z_n[nz - 1] - zbar_n[nz]);
            vd[nz] = wup * SG(nz - 1) * d->Ki[LAY(n, nz - 1)];
            vd[nz] = vd[nz] + (zbar_n[nz] - z_n[nz]) * SG(nz) * d->Ki[LAY(n, nz)];
            vd[nz] = vd[nz] / (z_n[nz - 1] - z_n[nz]) * m->area[INT(n, nz)];
        }
        for (int nz = ul1; nz <= nl1; ++nz)
            del_ttf[LAY(n, nz)] = del_ttf[LAY(n, nz)] + (vd[nz] - vd[nz + 1]) * dt / m->areasvol[INT(n, nz)];
    }
#undef SG
    free(txyn); free(vd); free(zbar_n); free(z_n);
}

/* ------------------------------------------------------------------------ */
/* bc_surface: oce_ale_tracer.F90:1828-2108 (non-RECOM IDs)                  */
/* ------------------------------------------------------------------------ */
int fo_bc_surface(const fo_mesh *m, int id, double dt, const fo_dyn *d, int n, int nzmin, double sval, double *out)
{
    switch (id) {
    case 1:
        *out = -dt * (d->heat_flux[n] / VCPW + sval * d->water_flux[n] * d->is_nonlinfs);
        return 0;
    case 2:
        *out = dt * (d->virtual_salt[n] + d->relax_salt[n] + d->real_salt_flux[n] * d->is_nonlinfs);
        return 0;
    case 100: case 301: case 302: case 303:
        *out = 0.0;
        return 0;
    case 101:   /* :2059-2060 */
        if (!d->prec_rain) return -1;
        *out = dt * d->prec_rain[n];
        return 0;
    case 102: case 103: case 104:   /* :2062-2067, wiso_flux_oce(n, ID-101) */
        if (!d->wiso_flux_oce) return -1;
        *out = dt * d->wiso_flux_oce[(size_t)(g_mut == 47 ? 104 - id : id - 102) * m->nod2D + n];
        return 0;
    case 304:   /* :2074-2079, nzmin is the level being filled */
        if (d->hosing_depth) {
            if (!d->hosing_flux3D) return -1;
            const int kk = (g_mut == 48 && nzmin < m->nl - 1) ? nzmin + 1 : nzmin;
            *out = dt * d->hosing_flux3D[(size_t)n * (m->nl - 1) + (kk - 1)];
        } else {
            if (!d->hosing_flux) return -1;
            *out = dt * d->hosing_flux[n];
        }
        return 0;
    case 501:   /* ice-shelf water from basal melting: under a cavity only (:2080-2085) */
        *out = (nzmin == 1 && g_mut != 46) ? 0.0 : dt * d->water_flux[n] * (sval - 1.0);
        return 0;
    default:
        return -1;   /* par_ex: invalid ID (oce_ale_tracer.F90:2091-2098) */
    }
}

/* ------------------------------------------------------------------------ */
/* diff_ver_part_impl_ale: oce_ale_tracer.F90:907-1435                       */
/* (Redi off -> isredi = 0; FCT or no w-split -> do_wimpl false; no KPP       */
/*  non-local flux, no shortwave penetration, no icebergs, no hosing)        */
/* ------------------------------------------------------------------------ */
int fo_diff_ver_part_impl_ale(const fo_mesh *m, const fo_dyn *d, double dt, fo_tracer *t)
{
    const int L = m->nl - 1, NL = m->nl;
    /* bc_surface fails for every column or for none (an unknown ID, a flux
       array that is not set): checked once, before any column changes */
    double bc0;
    if (m->nod2D > 0 && fo_bc_surface(m, t->ID, dt, d, 0, m->ulevels_nod2D[0], 0.0, &bc0) != 0) return -1;
#ifdef _OPENMP
#pragma omp parallel
#endif
    {
    double *a = (double *)calloc(NL, sizeof(double)), *b = (double *)calloc(NL, sizeof(double));
    double *c = (double *)calloc(NL, sizeof(double)), *tr = (double *)calloc(NL, sizeof(double));
    double *cp = (double *)calloc(NL, sizeof(double)), *tp = (double *)calloc(NL, sizeof(double));
    double *zbar_n = (double *)calloc(NL, sizeof(double)), *Z_n = (double *)calloc(NL, sizeof(double));
#ifdef _OPENMP
#pragma omp for schedule(static)
#endif
    for (int n = 0; n < m->nod2D; ++n) {
        const double *hn = m->hnode_new + LAY(n, 1);
        const double *ar = m->area + INT(n, 1);
        const double *av = m->areasvol + INT(n, 1);
        const double *kv = d->Kv + INT(n, 1);
        double *T = t->values + LAY(n, 1);
#define I(k) ((k) - 1)
        for (int i = 0; i < NL; ++i) { a[i] = b[i] = c[i] = tr[i] = tp[i] = cp[i] = 0.0; zbar_n[i] = Z_n[i] = 0.0; }
        int nzmax = m->nlevels_nod2D[n], nzmin = m->ulevels_nod2D[n];
        zbar_n[I(nzmax)] = m->zbar_n_bot[n];
        Z_n[I(nzmax - 1)] = zbar_n[I(nzmax)] + hn[I(nzmax - 1)] / 2.0;
        for (int nz = nzmax - 1; nz >= nzmin + 1; --nz) {
            zbar_n[I(nz)] = zbar_n[I(nz + 1)] + hn[I(nz)];
            Z_n[I(nz - 1)] = zbar_n[I(nz)] + hn[I(nz - 1)] / 2.0;
        }
        zbar_n[I(nzmin)] = zbar_n[I(nzmin + 1)] + hn[I(nzmin)];
        /* implicit part of the split vertical velocity (oce_ale_tracer.F90:961,
           1118-1129, 1158-1172, 1194-1202): FCT tracers take it in the LO step */
        const int wimpl = d->use_wsplit && t->lim != FO_LIM_FCT;
        const double *wi = d->w_i + INT(n, 1);
        const double ws = (g_mut == 28) ? -1.0 : 1.0;
        int nz = nzmin;
        double zinv2 = 1.0 / (Z_n[I(nz)] - Z_n[I(nz + 1)]);
        double zinv = 1.0 * dt;
        int ic = (g_mut == 12) ? nz : nz + 1;
        /* Redi: K33 = Kv + Ki S3^2 interpolated to the interfaces
           (oce_ale_tracer.F90:1105-1108, 1138-1144, 1181-1184) */
        const int redi = d->redi && d->Ki && d->slope_tapered;
#define S3(k) d->slope_tapered[LAY(n, (k)) * 3 + 2]
#define KI(k) d->Ki[LAY(n, (k))]
        a[I(nz)] = 0.0;
        if (redi) {
            const double Ty1 = (zbar_n[I(nz + 1)] - Z_n[I(nz + 1)]) * zinv2 * (S3(nz) * S3(nz)) * KI(nz)
                               + (Z_n[I(nz)] - zbar_n[I(nz + 1)]) * zinv2 * (S3(nz + 1) * S3(nz + 1)) * KI(nz + 1);
            c[I(nz)] = -(kv[I(nz + 1)] + Ty1) * zinv2 * zinv * ar[I(ic)] / av[I(nz)];
        } else {
            c[I(nz)] = -(kv[I(nz + 1)]) * zinv2 * zinv * ar[I(ic)] / av[I(nz)];
        }
        b[I(nz)] = -c[I(nz)] + hn[I(nz)];
        if (wimpl) {
            double v_adv = zinv * (ar[I(nz)] / av[I(nz)]);
            b[I(nz)] = b[I(nz)] + ws * wi[I(nz)] * v_adv;
            v_adv = zinv * ar[I(nz + 1)] / av[I(nz)];
            b[I(nz)] = b[I(nz)] - dmin(0.0, wi[I(nz + 1)]) * v_adv;
            c[I(nz)] = c[I(nz)] - dmax(0.0, wi[I(nz + 1)]) * v_adv;
        }
        double zinv1 = zinv2;
        for (nz = nzmin + 1; nz <= nzmax - 2; ++nz) {
            zinv2 = 1.0 / (Z_n[I(nz)] - Z_n[I(nz + 1)]);
            ic = (g_mut == 12) ? nz : nz + 1;
            if (redi) {
                const double zt = (g_mut == 33) ? zinv2 : zinv1;
                const double Ty = (zbar_n[I(nz)] - Z_n[I(nz)]) * zt * (S3(nz - 1) * S3(nz - 1)) * KI(nz - 1)
                                  + (Z_n[I(nz - 1)] - zbar_n[I(nz)]) * zt * (S3(nz) * S3(nz)) * KI(nz);
                const double Ty1 = (zbar_n[I(nz + 1)] - Z_n[I(nz + 1)]) * zinv2 * (S3(nz) * S3(nz)) * KI(nz)
                                   + (Z_n[I(nz)] - zbar_n[I(nz + 1)]) * zinv2 * (S3(nz + 1) * S3(nz + 1)) * KI(nz + 1);
                a[I(nz)] = -(kv[I(nz)] + Ty) * zinv1 * zinv * (ar[I(nz)] / av[I(nz)]);
                c[I(nz)] = -(kv[I(nz + 1)] + Ty1) * zinv2 * zinv * ar[I(ic)] / av[I(nz)];
            } else {
                a[I(nz)] = -(kv[I(nz)]) * zinv1 * zinv * (ar[I(nz)] / av[I(nz)]);
                c[I(nz)] = -(kv[I(nz + 1)]) * zinv2 * zinv * ar[I(ic)] / av[I(nz)];
            }
            b[I(nz)] = -a[I(nz)] - c[I(nz)] + hn[I(nz)];
            zinv1 = zinv2;
            if (wimpl) {
                double v_adv = zinv * (ar[I(nz)] / av[I(nz)]);
                a[I(nz)] = a[I(nz)] + dmin(0.0, wi[I(nz)]) * v_adv;
                b[I(nz)] = b[I(nz)] + dmax(0.0, wi[I(nz)]) * v_adv;
                v_adv = zinv * ar[I(nz + 1)] / av[I(nz)];
                b[I(nz)] = b[I(nz)] - dmin(0.0, wi[I(nz + 1)]) * v_adv;
                c[I(nz)] = c[I(nz)] - dmax(0.0, wi[I(nz + 1)]) * v_adv;
            }
        }
        nz = nzmax - 1;
        zinv = 1.0 * dt;
        if (redi) {
            const double Ty = (zbar_n[I(nz)] - Z_n[I(nz)]) * zinv1 * (S3(nz - 1) * S3(nz - 1)) * KI(nz - 1)
                              + (Z_n[I(nz - 1)] - zbar_n[I(nz)]) * zinv1 * (S3(nz) * S3(nz)) * KI(nz);
            a[I(nz)] = -(kv[I(nz)] + Ty) * zinv1 * zinv * (ar[I(nz)] / av[I(nz)]);
        } else {
            a[I(nz)] = -(kv[I(nz)]) * zinv1 * zinv * (ar[I(nz)] / av[I(nz)]);
        }
#undef S3
#undef KI
        c[I(nz)] = 0.0;
        b[I(nz)] = -a[I(nz)] + hn[I(nz)];
        if (wimpl) {
            const double v_adv = zinv * (ar[I(nz)] / av[I(nz)]);
            a[I(nz)] = a[I(nz)] + dmin(0.0, wi[I(nz)]) * v_adv;
            b[I(nz)] = b[I(nz)] + dmax(0.0, wi[I(nz)]) * v_adv;
        }

        nz = nzmin;
        double dz = hn[I(nz)];
        tr[I(nz)] = -(b[I(nz)] - dz) * T[I(nz)] - c[I(nz)] * T[I(nz + 1)];
        for (nz = nzmin + 1; nz <= nzmax - 2; ++nz) {
            dz = hn[I(nz)];
            tr[I(nz)] = -a[I(nz)] * T[I(nz - 1)] - (b[I(nz)] - dz) * T[I(nz)] - c[I(nz)] * T[I(nz + 1)];
        }
        nz = nzmax - 1;
        dz = hn[I(nz)];
        tr[I(nz)] = -a[I(nz)] * T[I(nz - 1)] - (b[I(nz)] - dz) * T[I(nz)];

        double bc = 0.0;
        if (t->ID == 304 && d->hosing_depth) {
            /* the hosing tracer in depth mode takes its source at every level
               nzmin..nzmax (oce_ale_tracer.F90:1374-1380); row nzmax is not
               part of the system, so the wet layers nzmin..nzmax-1 receive it */
            for (nz = nzmin; nz <= nzmax - 1; ++nz) {
                (void)fo_bc_surface(m, t->ID, dt, d, n, nz, T[I(nz)], &bc);
                tr[I(nz)] = tr[I(nz)] + bc;
            }
        } else {
            (void)fo_bc_surface(m, t->ID, dt, d, n, nzmin, T[I(nzmin)], &bc);
            if (g_mut == 41) tr[I(1)] = tr[I(1)] + bc;
            else tr[I(nzmin)] = tr[I(nzmin)] + bc;
        }

        cp[I(nzmin)] = c[I(nzmin)] / b[I(nzmin)];
        tp[I(nzmin)] = tr[I(nzmin)] / b[I(nzmin)];
        for (nz = nzmin + 1; nz <= nzmax - 1; ++nz) {
            double mm = b[I(nz)] - cp[I(nz - 1)] * a[I(nz)];
            cp[I(nz)] = c[I(nz)] / mm;
            tp[I(nz)] = (tr[I(nz)] - tp[I(nz - 1)] * a[I(nz)]) / mm;
        }
        tr[I(nzmax - 1)] = tp[I(nzmax - 1)];
        for (nz = nzmax - 2; nz >= nzmin; --nz) {
            if (g_mut == 13) tr[I(nz)] = tp[I(nz)] + cp[I(nz)] * tr[I(nz + 1)];
            else tr[I(nz)] = tp[I(nz)] - cp[I(nz)] * tr[I(nz + 1)];
        }
        for (nz = nzmin; nz <= nzmax - 1; ++nz) T[I(nz)] = T[I(nz)] + tr[I(nz)];
#undef I
    }
    free(a); free(b); free(c); free(tr); free(cp); free(tp); free(zbar_n); free(Z_n);
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* one tracer step for all tracers: solve_tracers_ale, oce_ale_tracer.F90:   */
/* 280-385 (per tracer) and the salinity clamp at :502-514                   */
/* ------------------------------------------------------------------------ */
int fo_solve_tracers_ale(const fo_mesh *m, const fo_dyn *d, double dt, fo_tracer *t, int ntr, fo_work *w)
{
    const int L = m->nl - 1;
    for (int i = 0; i < ntr; ++i) {
        int rc = fo_init_tracers_AB(m, &t[i], w);
        if (rc) return rc;
        rc = fo_do_oce_adv_tra(m, d, dt, &t[i], w);
        if (rc) return rc;
        fo_diff_tracers_ale_tstar(m, d, dt, &t[i], w);
        if (t[i].i_vert_diff) {
            rc = fo_diff_ver_part_impl_ale(m, d, dt, &t[i]);
            if (rc) return rc;
        }
    }
    if (ntr >= 2) {
        double *s = t[1].values;
        OMP_FOR
        for (int n = 0; n < m->nod2D; ++n)
            for (int nz = m->ulevels_nod2D[n]; nz <= m->nlevels_nod2D[n] - 1; ++nz) {
                if (s[LAY(n, nz)] > 45.0) s[LAY(n, nz)] = 45.0;
                if (s[LAY(n, nz)] < 3.0) s[LAY(n, nz)] = 3.0;
            }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* init_stiff_mat_ale: oce_ale.F90:1393-1582 (one partition, local numbering)*/
/* ------------------------------------------------------------------------ */
int fo_init_stiff_mat_ale(const fo_mesh *m, double dt, int *rowptr, int *colind, double *values, int max_nnz)
{
    const int N2 = m->nod2D;
    int *n_num = (int *)calloc(N2, sizeof(int));
    int *n_pos = (int *)calloc((size_t)N2 * 16, sizeof(int));
    for (int n = 0; n < N2; ++n) { n_num[n] = 1; n_pos[(size_t)n * 16] = n; }
    for (int e = 0; e < m->edge2D; ++e) {
        int n1 = m->edges[2 * e], n2 = m->edges[2 * e + 1];
        if (n_num[n1] >= 16 || n_num[n2] >= 16) { free(n_num); free(n_pos); return -1; }
        n_pos[(size_t)n1 * 16 + n_num[n1]] = n2; n_num[n1]++;
        n_pos[(size_t)n2 * 16 + n_num[n2]] = n1; n_num[n2]++;
    }
    rowptr[0] = 0;
    for (int n = 0; n < N2; ++n) rowptr[n + 1] = rowptr[n] + n_num[n];
    if (rowptr[N2] > max_nnz) { free(n_num); free(n_pos); return -2; }
    for (int n = 0; n < N2; ++n)
        for (int j = 0; j < n_num[n]; ++j) colind[rowptr[n] + j] = n_pos[(size_t)n * 16 + j];
    for (int i = 0; i < rowptr[N2]; ++i) values[i] = 0.0;
    /* position map: n_num(col) = position within the current row */
    int *pos = (int *)malloc(sizeof(int) * (size_t)N2);
    for (int n = 0; n < N2; ++n) pos[n] = -1;
    const double factor = G_ACC * dt * 1.0 * 1.0;   /* g*dt*alpha*theta */
    for (int ed = 0; ed < m->edge2D; ++ed) {
        for (int i = 0; i < 2; ++i) {
            int el = m->edge_tri[2 * ed + i];
            if (el < 0) continue;
            const int *en = m->elem2D_nodes + 3 * (size_t)el;
            const double *gs = m->gradient_sca + 6 * (size_t)el;
            double hh = m->zbar_e_bot[el] - m->zbar_e_srf[el];
            double cx = m->edge_cross_dxdy[4 * ed + 2 * i];       /* edge_cross_dxdy(2*i-1) */
            double cy = m->edge_cross_dxdy[4 * ed + 2 * i + 1];   /* edge_cross_dxdy(2*i)   */
            double fy[3];
            for (int j = 0; j < 3; ++j) fy[j] = hh * (gs[j] * cy - gs[3 + j] * cx);
            if (i == 1) for (int j = 0; j < 3; ++j) fy[j] = -fy[j];
            for (int side = 0; side < 2; ++side) {
                int row = m->edges[2 * ed + side];
                for (int p = rowptr[row]; p < rowptr[row + 1]; ++p) pos[colind[p]] = p;
                for (int j = 0; j < 3; ++j) {
                    int p = pos[en[j]];
                    if (side == 0) values[p] = values[p] + fy[j] * factor;
                    else values[p] = values[p] - fy[j] * factor;
                }
            }
        }
    }
    for (int row = 0; row < N2; ++row) {
        if (m->ulevels_nod2D[row] > 1 && g_mut != 38) continue;
        int off = rowptr[row];
        double av = m->areasvol[(size_t)row * m->nl + (m->ulevels_nod2D[row] - 1)];
        if (g_mut == 14) values[off] = values[off] + av * dt;
        else values[off] = values[off] + av / dt;
    }
    free(pos); free(n_num); free(n_pos);
    return rowptr[N2];
}

/* ------------------------------------------------------------------------ */
/* ssh_solve_preconditioner: solver.F90:31-117                               */
/* variant 0/1 as in the reference; variant 2 = diagonal (Jacobi) only       */
/* ------------------------------------------------------------------------ */
int fo_ssh_solve_preconditioner(int n, const int *rowptr, const int *colind, const double *values,
                                int variant, double *pr)
{
    double *diag = (double *)malloc(sizeof(double) * (size_t)n);
    for (int r = 0; r < n; ++r) diag[r] = values[rowptr[r]];
    for (int r = 0; r < n; ++r) {
        int off = rowptr[r];
        int nend = rowptr[r + 1] - rowptr[r];
        pr[off] = 1.0 / values[off];
        for (int j = 1; j < nend; ++j) {
            int node = colind[off + j];
            if (variant == 0)
                pr[off + j] = -0.5 * (values[off + j] / values[off]) / (values[off] + diag[node]);
            else if (variant == 1) {
                double v = -values[off + j] / (values[off] * diag[node]);
                pr[off + j] = (g_mut == 16) ? -v : v;
            } else
                pr[off + j] = 0.0;
        }
    }
    free(diag);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* ssh_solve_cg: solver.F90:120-385 (one partition; serial sums)             */
/* ------------------------------------------------------------------------ */
static double row_dot(const int *rowptr, const int *colind, const double *v, const double *x, int r)
{
    /* Fortran sum(values(a:b) * x(cind(a:b))): left-to-right */
    double s = 0.0;
    for (int p = rowptr[r]; p < rowptr[r + 1]; ++p) s = s + v[p] * x[colind[p]];
    return s;
}

int fo_ssh_solve_cg(int n, const int *rowptr, const int *colind, const double *values,
                    const double *pr, double *x, const double *rhs, double soltol,
                    int maxiter, fo_cg_stats *st)
{
    double *rr = (double *)calloc(n, sizeof(double)), *zz = (double *)calloc(n, sizeof(double));
    double *pp = (double *)calloc(n, sizeof(double)), *App = (double *)calloc(n, sizeof(double));
    double s_old = 0.0, s_aux, al, be, sprod[2] = {0.0, 0.0};
    for (int r = 0; r < n; ++r) s_old = s_old + rhs[r] * rhs[r];
    double rtol = soltol * sqrt(s_old / (double)n);
    for (int r = 0; r < n; ++r) rr[r] = rhs[r] - row_dot(rowptr, colind, values, x, r);
    for (int r = 0; r < n; ++r) { zz[r] = row_dot(rowptr, colind, pr, rr, r); pp[r] = zz[r]; }
    s_old = 0.0;
    for (int r = 0; r < n; ++r) s_old = s_old + rr[r] * zz[r];
    int converged = 0, iter;
    st->nbreakdown = 0;
    st->nnegrz = 0;
    for (iter = 1; iter <= maxiter; ++iter) {
        for (int r = 0; r < n; ++r) App[r] = row_dot(rowptr, colind, values, pp, r);
        s_aux = 0.0;
        for (int r = 0; r < n; ++r) s_aux = s_aux + pp[r] * App[r];
        if (s_aux <= 0.0 || s_aux != s_aux) { st->nbreakdown++; break; }
        al = s_old / s_aux;
        for (int r = 0; r < n; ++r) { x[r] = x[r] + al * pp[r]; rr[r] = rr[r] - al * App[r]; }
        for (int r = 0; r < n; ++r) zz[r] = row_dot(rowptr, colind, pr, rr, r);
        sprod[0] = 0.0; sprod[1] = 0.0;
        for (int r = 0; r < n; ++r) { sprod[0] = sprod[0] + rr[r] * zz[r]; sprod[1] = sprod[1] + rr[r] * rr[r]; }
        if (sqrt(sprod[1] / (double)n) < rtol) { converged = 1; break; }
        if (sprod[0] < 0.0) st->nnegrz++;
        if (s_old == 0.0) { st->nbreakdown++; break; }
        be = (g_mut == 15) ? s_old / sprod[0] : sprod[0] / s_old;
        s_old = sprod[0];
        for (int r = 0; r < n; ++r) pp[r] = zz[r] + be * pp[r];
    }
    st->iters_last = iter < maxiter ? iter : maxiter;
    st->converged = converged;
    st->resid_last = sqrt(sprod[1] / (double)n);
    st->rtol_last = rtol;
    free(rr); free(zz); free(pp); free(App);
    return converged ? 0 : 1;
}

/* ------------------------------------------------------------------------ */
/* vert_vel_ale: oce_ale.F90:2107-2668, which_ALE = 'linfs' (0), 'zlevel'  */
/* (1) or 'zstar' (2); then compute_CFLz (oce_ale.F90:2910-2997) and        */
/* compute_Wvel_split (oce_ale.F90:3003-3049).                              */
/* Outputs (N2*NL): w, w_e, w_i, cfl_z. hnode_new (N2*L) holds the current  */
/* thicknesses on entry and is updated in place under zstar.               */
/* ------------------------------------------------------------------------ */
int fo_vert_vel_ale(const fo_mesh *m, const double *uv, double dt, int which_ale, const double *hbar,
                    const double *hbar_old, const double *water_flux, int use_wsplit, double wsplit_maxcfl,
                    double *w, double *w_e, double *w_i, double *cfl_z, double *hnode_new, double *cflmax,
                    const double *zbar, double min_hnode, int lzstar_lev)
{
    const int L = m->nl - 1, NL = m->nl, N2 = m->nod2D;
    if (which_ale != 0 && which_ale != 1 && which_ale != 2) return -1;
    if (which_ale != 0 && (!hbar || !hbar_old || !water_flux)) return -2;
    if (which_ale == 1 && (!zbar || lzstar_lev < 1 || lzstar_lev > L)) return -3;
    /* :2164-2171 */
    memset(w, 0, sizeof(double) * (size_t)N2 * NL);
    /* thickness divergence, edges ascending (:2175-2265): each element of
       the edge adds its layer transports to enodes(1) and subtracts them
       from enodes(2), el(1) before el(2) */
    for (int ed = 0; ed < m->edge2D; ++ed) {
        const int n1 = m->edges[2 * ed], n2 = m->edges[2 * ed + 1];
        for (int s = 0; s < 2; ++s) {
            const int el = m->edge_tri[2 * ed + s];
            if (el < 0) continue;
            const double dX = m->edge_cross_dxdy[4 * ed + 2 * s], dY = m->edge_cross_dxdy[4 * ed + 2 * s + 1];
            const int nzmin = m->ulevels[el], nzmax = m->nlevels[el] - 1;
            for (int nz = nzmin; nz <= nzmax; ++nz) {
                const double *u = uv + ((size_t)el * L + nz - 1) * 2;
                double c1 = (u[1] * dX - u[0] * dY) * m->helem[LAY(el, nz)];
                if (s == 1 && g_mut != 19) c1 = -c1;
                w[INT(n1, nz)] = w[INT(n1, nz)] + c1;
                w[INT(n2, nz)] = w[INT(n2, nz)] - c1;
            }
        }
    }
    /* cumulative sum from the bottom (:2277-2299), then divide by the
       interface area (:2301-2320) */
    for (int n = 0; n < N2; ++n) {
        const int nzmin = m->ulevels_nod2D[n], nzmax = m->nlevels_nod2D[n] - 1;
        if (g_mut == 20) {
            for (int nz = nzmin + 1; nz <= nzmax; ++nz) w[INT(n, nz)] = w[INT(n, nz)] + w[INT(n, nz - 1)];
        } else {
            for (int nz = nzmax; nz >= nzmin; --nz) w[INT(n, nz)] = w[INT(n, nz)] + w[INT(n, nz + 1)];
        }
        for (int nz = nzmin; nz <= nzmax; ++nz) w[INT(n, nz)] = w[INT(n, nz)] / m->area[INT(n, nz)];
    }
    /* zlevel (:2336-2420): the ssh change goes into the surface layer; when
       that would thin it below min_hnode of its nominal thickness, it is
       spread over the top lzstar_lev layers (local zstar), skipping layers
       with CFL_z >= 0.95 from the previous step; a rising surface first
       restores thinned layers. Cavity columns stay linfs. */
    if (which_ale == 1) {
        const int lz = lzstar_lev;
        double *mx = (double *)calloc(lz + 1, sizeof(double)), *cum = (double *)calloc(lz + 1, sizeof(double));
        double *dist = (double *)calloc(lz + 1, sizeof(double));
        const double mh = (g_mut == 37) ? 0.0 : min_hnode;
        for (int n = 0; n < N2; ++n) {
            const int nzmin = m->ulevels_nod2D[n];
            int nzmax = m->nlevels_nod2D_min[n] - 1;
            if (nzmin != 1) continue;
#define HO(k) m->hnode[LAY(n, (k))]
            const double dtot = hbar[n] - hbar_old[n];
            if (dtot < 0.0 && HO(nzmin) + dtot <= (zbar[nzmin - 1] - zbar[nzmin]) * mh) {
                for (int j = 1; j <= lz; ++j) mx[j] = (zbar[nzmin + j - 2] - zbar[nzmin + j - 1]) * mh - HO(nzmin + j - 1);
                for (int j = 1; j <= lz; ++j) if (mx[j] >= 0.0) mx[j] = 0.0;
                for (int j = 1; j <= lz; ++j) if (cfl_z[INT(n, nzmin + j - 1)] >= 0.95) mx[j] = 0.0;
                cum[1] = mx[1];
                for (int j = 2; j <= lz; ++j) cum[j] = mx[j] + mx[j - 1];
                int nz = lz;   /* minval of an empty pack is huge(); min with lzstar_lev */
                for (int j = lz; j >= 1; --j) if (cum[j] < dtot) nz = j;
                for (int j = 1; j <= lz; ++j) dist[j] = 0.0;
                double rest = dtot;
                nzmax = imin(nz, nzmax - 1);
                for (int j = 1; j <= nzmax; ++j) {
                    dist[j] = dmax(rest, mx[j]);
                    rest = rest - dist[j];
                    rest = dmin(0.0, rest);
                }
                double dint = 0.0;
                for (int j = nzmax; j >= 1; --j) {
                    dint = dint + dist[j];
                    w[INT(n, nzmin + j - 1)] = w[INT(n, nzmin + j - 1)] - ((g_mut == 36) ? dist[j] : dint) / dt;
                    hnode_new[LAY(n, nzmin + j - 1)] = HO(nzmin + j - 1) + dist[j];
                }
            } else {
                int thin = 0;
                for (int j = 2; j <= lz; ++j) thin |= HO(nzmin + j - 1) != (zbar[nzmin + j - 2] - zbar[nzmin + j - 1]);
                if (dtot > 0.0 && thin) {
                    for (int j = 1; j <= lz; ++j) mx[j] = (zbar[nzmin + j - 2] - zbar[nzmin + j - 1]) - HO(nzmin + j - 1);
                    mx[1] = 1000.0;
                    int nz = 0;
                    for (int j = 1; j <= lz; ++j)
                        if (HO(nzmin + j - 1) != (zbar[nzmin + j - 2] - zbar[nzmin + j - 1])) nz = j;
                    nzmax = imin(nz, nzmax - 1);
                    double rest = dtot, dint = 0.0;
                    for (int j = 1; j <= lz; ++j) dist[j] = 0.0;
                    for (int j = nzmax; j >= 1; --j) {
                        dist[j] = dmin(rest, mx[j]);
                        rest = rest - dist[j];
                        rest = dmax(0.0, rest);
                        dint = dint + dist[j];
                        w[INT(n, nzmin + j - 1)] = w[INT(n, nzmin + j - 1)] - ((g_mut == 36) ? dist[j] : dint) / dt;
                        hnode_new[LAY(n, nzmin + j - 1)] = HO(nzmin + j - 1) + dist[j];
                    }
                } else {
                    w[INT(n, nzmin)] = w[INT(n, nzmin)] - dtot / dt;
                    hnode_new[LAY(n, nzmin)] = HO(nzmin) + dtot;
                }
            }
            w[INT(n, nzmin)] = w[INT(n, nzmin)] - water_flux[n];
#undef HO
        }
        free(mx); free(cum); free(dist);
    }
    /* zstar: distribute hbar - hbar_old over the layers above
       nlevels_nod2D_min (:2539-2600); cavity columns stay linfs */
    if (which_ale == 2) {
        for (int n = 0; n < N2; ++n) {
            const int nzmin = m->ulevels_nod2D[n], nzmax = m->nlevels_nod2D_min[n] - 1;
            if (nzmin != 1 && g_mut != 39) continue;
            const double dd1 = m->zbar_3d_n[INT(n, nzmax)];
            double dd = m->zbar_3d_n[INT(n, nzmin)] - dd1;
            dd = (hbar[n] - hbar_old[n]) / dd;
            const double dddt = dd / dt;
            for (int nz = nzmin; nz <= nzmax - 1; ++nz) {
                w[INT(n, nz)] = w[INT(n, nz)] - (m->zbar_3d_n[INT(n, nz)] - dd1) * dddt;
                hnode_new[LAY(n, nz)] = m->hnode[LAY(n, nz)] + (m->zbar_3d_n[INT(n, nz)] - m->zbar_3d_n[INT(n, nz + 1)]) * dd;
            }
            w[INT(n, nzmin)] = w[INT(n, nzmin)] - water_flux[n];
        }
    }
    /* compute_CFLz (:2938-2966): interface nz collects |W| dt / h of the
       layers below (c1) and above (c2) it */
    memset(cfl_z, 0, sizeof(double) * (size_t)N2 * NL);
    double cmax = 0.0;
    for (int n = 0; n < N2; ++n) {
        const int nzmin = m->ulevels_nod2D[n], nzmax = m->nlevels_nod2D[n] - 1;
        for (int nz = nzmin; nz <= nzmax; ++nz) {
            const double c1 = fabs(w[INT(n, nz)] * dt / hnode_new[LAY(n, nz)]);
            const double c2 = fabs(w[INT(n, g_mut == 22 ? nz : nz + 1)] * dt / hnode_new[LAY(n, nz)]);
            cfl_z[INT(n, nz)] = cfl_z[INT(n, nz)] + c1;
            cfl_z[INT(n, nz + 1)] = c2;
        }
        for (int k = 1; k <= NL; ++k) cmax = dmax(cmax, cfl_z[INT(n, k)]);
    }
    if (cflmax) *cflmax = cmax;
    /* compute_Wvel_split (:3035-3047) */
    for (int n = 0; n < N2; ++n) {
        const int nzmin = m->ulevels_nod2D[n], nzmax = m->nlevels_nod2D[n];
        for (int k = 1; k <= NL; ++k) { w_e[INT(n, k)] = 0.0; w_i[INT(n, k)] = 0.0; }
        for (int nz = nzmin; nz <= nzmax; ++nz) {
            const size_t i = INT(n, nz);
            w_e[i] = w[i];
            w_i[i] = 0.0;
            if (use_wsplit && (cfl_z[i] > wsplit_maxcfl)) {
                const double dd = dmax((cfl_z[i] - wsplit_maxcfl), 0.0) / dmax(wsplit_maxcfl, 1.e-12);
                w_e[i] = (1.0 / (1.0 + dd)) * w[i];
                w_i[i] = (dd / (1.0 + dd)) * w[i];
            }
        }
    }
    return 0;
}

/* ------------------------------------------------------------------------ */
/* compute_hbar_ale: oce_ale.F90:1956-2091. Column-integrated transport     */
/* divergence per node (edge scatter in edge order), the freshwater flux   */
/* off linfs, the new hbar and the element thickness change dhe.            */
/* ------------------------------------------------------------------------ */
int fo_compute_hbar_ale(const fo_mesh *m, const double *uv, const double *water_flux, double dt, int which_ale,
                        double *hbar, double *hbar_old, double *ssh_rhs_old, double *dhe)
{
    const int L = m->nl - 1, NL = m->nl, N2 = m->nod2D;
    for (int n = 0; n < N2; ++n) ssh_rhs_old[n] = 0.0;
    for (int ed = 0; ed < m->edge2D; ++ed) {
        const int n1 = m->edges[2 * ed], n2 = m->edges[2 * ed + 1];
        const int el1 = m->edge_tri[2 * ed], el2 = m->edge_tri[2 * ed + 1];
        double c1 = 0.0, c2 = 0.0;
        {
            const double dX = m->edge_cross_dxdy[4 * ed], dY = m->edge_cross_dxdy[4 * ed + 1];
            for (int nz = m->ulevels[el1]; nz <= m->nlevels[el1] - 1; ++nz) {
                const size_t i = ((size_t)el1 * L + nz - 1) * 2;
                c1 = c1 + (uv[i + 1] * dX - uv[i] * dY) * m->helem[LAY(el1, nz)];
            }
        }
        if (el2 >= 0) {
            const double dX = m->edge_cross_dxdy[4 * ed + 2], dY = m->edge_cross_dxdy[4 * ed + 3];
            for (int nz = m->ulevels[el2]; nz <= m->nlevels[el2] - 1; ++nz) {
                const size_t i = ((size_t)el2 * L + nz - 1) * 2;
                c2 = c2 - (uv[i + 1] * dX - uv[i] * dY) * m->helem[LAY(el2, nz)];
            }
        }
        ssh_rhs_old[n1] = ssh_rhs_old[n1] + (c1 + c2);
        ssh_rhs_old[n2] = ssh_rhs_old[n2] - (c1 + c2);
    }
    if (which_ale != 0)
        for (int n = 0; n < N2; ++n) {
            if (m->ulevels_nod2D[n] > 1) continue;
            ssh_rhs_old[n] = ssh_rhs_old[n] - water_flux[n] * m->areasvol[INT(n, m->ulevels_nod2D[n])];
        }
    for (int n = 0; n < N2; ++n) hbar_old[n] = hbar[n];
    for (int n = 0; n < N2; ++n) {
        if (m->ulevels_nod2D[n] > 1) continue;   /* cavity node: hbar = hbar_old */
        const double av = (g_mut == 34) ? 1.0 : m->areasvol[INT(n, m->ulevels_nod2D[n])];
        hbar[n] = hbar_old[n] + ssh_rhs_old[n] * dt / av;
    }
    for (int el = 0; el < m->elem2D; ++el) {
        const int *en = m->elem2D_nodes + 3 * (size_t)el;
        if (m->ulevels[el] > 1) {
            dhe[el] = 0.0;
        } else {
            const double d3 = (g_mut == 35) ? 0.0 : (hbar[en[2]] - hbar_old[en[2]]);
            dhe[el] = ((hbar[en[0]] - hbar_old[en[0]]) + (hbar[en[1]] - hbar_old[en[1]]) + d3) / 3.0;
        }
    }
    (void)NL;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* compute_ssh_rhs_ale: oce_ale.F90:1803-1938. Depth-integrated transport of */
/* u + uv_rhs through the two dual faces of each edge, summed over levels    */
/* first (:1844-1880) and then into the edge's nodes, edges ascending        */
/* (:1889-1894); then the free-surface terms (:1913-1934).                   */
/* which_ale 0 = linfs, otherwise the full free surface branch.              */
/* ------------------------------------------------------------------------ */
int fo_compute_ssh_rhs_ale(const fo_mesh *m, const double *uv, const double *uv_rhs, const double *ssh_rhs_old,
                           const double *water_flux, double alpha, int which_ale, int use_cavity_fw2press,
                           double *ssh_rhs)
{
    const int L = m->nl - 1, NL = m->nl, N2 = m->nod2D;
    for (int n = 0; n < N2; ++n) ssh_rhs[n] = 0.0;
    for (int ed = 0; ed < m->edge2D; ++ed) {
        const int n1 = m->edges[2 * ed], n2 = m->edges[2 * ed + 1];
        const int el1 = m->edge_tri[2 * ed], el2 = m->edge_tri[2 * ed + 1];
        double c1 = 0.0, c2 = 0.0;
        {
            const double dX = m->edge_cross_dxdy[4 * ed], dY = m->edge_cross_dxdy[4 * ed + 1];
            for (int nz = m->ulevels[el1]; nz <= m->nlevels[el1] - 1; ++nz) {
                const size_t i = ((size_t)el1 * L + nz - 1) * 2;
                const double h = g_mut == 21 ? 1.0 : m->helem[LAY(el1, nz)];
                c1 = c1 + alpha * ((uv[i + 1] + uv_rhs[i + 1]) * dX - (uv[i] + uv_rhs[i]) * dY) * h;
            }
        }
        if (el2 >= 0) {
            const double dX = m->edge_cross_dxdy[4 * ed + 2], dY = m->edge_cross_dxdy[4 * ed + 3];
            for (int nz = m->ulevels[el2]; nz <= m->nlevels[el2] - 1; ++nz) {
                const size_t i = ((size_t)el2 * L + nz - 1) * 2;
                c2 = c2 - alpha * ((uv[i + 1] + uv_rhs[i + 1]) * dX - (uv[i] + uv_rhs[i]) * dY) * m->helem[LAY(el2, nz)];
            }
        }
        ssh_rhs[n1] = ssh_rhs[n1] + (c1 + c2);
        ssh_rhs[n2] = ssh_rhs[n2] - (c1 + c2);
    }
    if (which_ale != 0) {
        for (int n = 0; n < N2; ++n) {
            const int nzmin = m->ulevels_nod2D[n];
            if (nzmin > 1) {
                if (use_cavity_fw2press) ssh_rhs[n] = ssh_rhs[n] - alpha * water_flux[n] * m->areasvol[INT(n, nzmin)];
                if (g_mut == 40) ssh_rhs[n] = ssh_rhs[n] + (1.0 - alpha) * ssh_rhs_old[n];
            } else {
                ssh_rhs[n] = ssh_rhs[n] - alpha * water_flux[n] * m->areasvol[INT(n, nzmin)] + (1.0 - alpha) * ssh_rhs_old[n];
            }
        }
    } else {
        for (int n = 0; n < N2; ++n) ssh_rhs[n] = ssh_rhs[n] + (1.0 - alpha) * ssh_rhs_old[n];
    }
    (void)NL;
    return 0;
}

/* ------------------------------------------------------------------------ */
/* update_stiff_mat_ale: oce_ale.F90:1683-1792 (full free surface). The     */
/* stiffness term of the actual ssh change: -dhe(elem) in place of the      */
/* element's column depth, added to the assembled values; per row the edges  */
/* come in ascending order and each edge's elements in order, as in         */
/* fo_init_stiff_mat_ale.                                                    */
/* ------------------------------------------------------------------------ */
int fo_update_stiff_mat_ale(const fo_mesh *m, double dt, const double *dhe, const int *rowptr, const int *colind,
                            double *values)
{
    const int N2 = m->nod2D;
    int *pos = (int *)malloc(sizeof(int) * (size_t)N2);
    if (!pos) return -1;
    for (int n = 0; n < N2; ++n) pos[n] = -1;
    const double factor = G_ACC * dt * 1.0 * 1.0;   /* g*dt*alpha*theta */
    for (int ed = 0; ed < m->edge2D; ++ed) {
        for (int side = 0; side < 2; ++side) {
            const int row = m->edges[2 * ed + side];
            for (int p = rowptr[row]; p < rowptr[row + 1]; ++p) pos[colind[p]] = p;
            for (int i = 0; i < 2; ++i) {
                const int el = m->edge_tri[2 * ed + i];
                if (el < 0) continue;
                const int *en = m->elem2D_nodes + 3 * (size_t)el;
                const double *gs = m->gradient_sca + 6 * (size_t)el;
                const double cx = m->edge_cross_dxdy[4 * ed + 2 * i];
                const double cy = m->edge_cross_dxdy[4 * ed + 2 * i + 1];
                const double hh = g_mut == 23 ? dhe[el] : -dhe[el];
                double fy[3];
                for (int j = 0; j < 3; ++j) fy[j] = hh * (gs[j] * cy - gs[3 + j] * cx);
                if (i == 1) for (int j = 0; j < 3; ++j) fy[j] = -fy[j];
                if (side == 1) for (int j = 0; j < 3; ++j) fy[j] = -fy[j];
                for (int j = 0; j < 3; ++j) {
                    const int p = pos[en[j]];
                    if (p < 0) { free(pos); return -2; }
                    values[p] = values[p] + fy[j] * factor;
                }
            }
        }
    }
    free(pos);
    return 0;
}
