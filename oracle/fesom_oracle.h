/*
 * CPU oracle for the FESOM2 tracer-transport step and SSH solve.
 *
 * TEST INFRASTRUCTURE ONLY. This is a plain-C restatement of the reference
 * algorithm, used by tests/, by __graft_entry__.smoke() and by bench.py's CPU
 * baseline leg as the checker. It is never part of the product path.
 *
 * Conventions (same memory image as the reference's Fortran arrays):
 *   layer arrays      a[n*L + (k-1)],       L  = nl-1, k = 1..nl-1
 *   interface arrays  a[n*NL + (k-1)],      NL = nl,   k = 1..nl
 *   uv                uv[(e*L + k-1)*2 + c] (c = 0: u, 1: v)
 *   tr_xy             [(e*L + k-1)*2 + c]
 *   edge_up_dn_grad   [(e*L + k-1)*4 + c]
 * Connectivity is 0-based, -1 marks a missing element. Level numbers keep the
 * reference's 1-based meaning (a node's layers are ulevels..nlevels-1).
 */
#ifndef FESOM_ORACLE_H
#define FESOM_ORACLE_H

#ifdef __cplusplus
extern "C" {
#endif

typedef struct {
    int nod2D, elem2D, edge2D, nl, kmax, cartesian;
    double cyclic_length;
    const double *coord_nod2D;        /* N2*2 */
    const int *elem2D_nodes;          /* F2*3 */
    const int *edges, *edge_tri;      /* E2*2 */
    const int *nod_in_elem2D;         /* N2*kmax */
    const int *nod_in_elem2D_num;     /* N2 */
    const int *nlevels, *ulevels;     /* F2 */
    const int *nlevels_nod2D, *ulevels_nod2D, *nlevels_nod2D_min, *ulevels_nod2D_max; /* N2 */
    const double *elem_area, *elem_cos, *edge_dxdy, *edge_cross_dxdy, *gradient_sca;
    const double *area, *areasvol;    /* N2*nl */
    const double *hnode, *hnode_new;  /* N2*L */
    const double *helem;              /* F2*L */
    const double *zbar_3d_n;          /* N2*nl */
    const double *Z_3d_n;             /* N2*L */
    const double *zbar_n_bot;         /* N2 */
    const double *zbar_e_bot, *zbar_e_srf; /* F2 */
    const int *edge_up_dn_tri;        /* E2*2, from fo_find_up_downwind_triangles */
    /* node -> edge lists (2*e + side, side 1 where the node is edges(2,e)),
     * edges ascending: the OpenMP build sums the edge->node scatters as
     * per-node gathers in this order, which is the serial scatter's order */
    const int *nod_edge_ptr;          /* N2+1 */
    const int *nod_edge_idx;          /* 2*E2 */
} fo_mesh;

typedef struct {
    const double *uv, *w, *w_e, *w_i, *Kv;
    const double *heat_flux, *water_flux, *virtual_salt, *relax_salt, *real_salt_flux;
    double is_nonlinfs;
    int use_wsplit;
    const double *Ki;             /* N2*L (o_ARRAYS Ki); NULL: Ki = 0, no horizontal diffusion */
    const double *slope_tapered;  /* N2*L*3, component fastest; read when redi */
    int redi;                     /* Redi (oce_modules.F90:75) */
    /* per-node surface fluxes of the tracer IDs 101-104 and 304 (bc_surface,
     * oce_ale_tracer.F90:2059-2079); NULL where no such tracer runs */
    const double *prec_rain;      /* N2 (g_forcing_arrays prec_rain), ID 101 */
    const double *wiso_flux_oce;  /* N2*3, node fastest (o_ARRAYS wiso_flux_oce(:,1..3)), IDs 102-104 */
    const double *hosing_flux;    /* N2 (o_ARRAYS hosing_flux), ID 304 */
    const double *hosing_flux3D;  /* N2*L, level fastest (o_ARRAYS hosing_flux3D), ID 304 in depth mode */
    int hosing_depth;             /* use_hosing .and. hosing_mode == 'depth' */
} fo_dyn;

enum { FO_HOR_MFCT = 0, FO_HOR_UPW1 = 1, FO_HOR_MUSCL = 2 };
enum { FO_VER_QR4C = 0, FO_VER_UPW1 = 1, FO_VER_CDIFF = 2, FO_VER_PPM = 3 };
enum { FO_LIM_FCT = 0, FO_LIM_NON = 1 };

typedef struct {
    int ID, AB_order, hor, ver, lim, i_vert_diff;
    double ph, pv;
    double *values;      /* N2*L */
    double *valuesold;   /* N2*L*(AB_order-1), AB index fastest */
    double *valuesAB;    /* N2*L */
} fo_tracer;

typedef struct {
    double *tr_xy;            /* F2*L*2 */
    double *edge_up_dn_grad;  /* E2*L*4 */
    double *fct_LO;           /* N2*L */
    double *adv_flux_hor;     /* E2*L */
    double *adv_flux_ver;     /* N2*NL */
    double *fct_ttf_max, *fct_ttf_min, *fct_plus, *fct_minus; /* N2*L */
    double *del_ttf, *del_ttf_advhoriz, *del_ttf_advvert;    /* N2*L */
    double *tr_z;             /* N2*NL */
} fo_work;

/* planted-mistake switch for the pinning tests (0 = faithful) */
void fo_set_mutation(int m);
int fo_get_mutation(void);

int fo_find_up_downwind_triangles(const fo_mesh *m, int *edge_up_dn_tri);
int fo_init_tracers_AB(const fo_mesh *m, fo_tracer *t, fo_work *w);
void fo_tracer_gradient_elements(const fo_mesh *m, const double *ttf, double *tr_xy);
void fo_fill_up_dn_grad(const fo_mesh *m, const double *tr_xy, double *grad);
void fo_adv_tra_hor_upw1(const fo_mesh *m, const double *vel, const double *ttf, double *flux, int init_zero);
void fo_adv_tra_hor_mfct(const fo_mesh *m, const double *vel, const double *ttf, double num_ord,
                         double *flux, const double *grad, int init_zero);
void fo_adv_tra_hor_muscl(const fo_mesh *m, const double *vel, const double *ttf, double num_ord,
                          double *flux, const double *grad, const int *nblay, int init_zero);
void fo_muscl_nboundary_lay(const fo_mesh *m, int *nblay);
void fo_adv_tra_ver_upw1(const fo_mesh *m, const double *w, const double *ttf, double *flux, int init_zero);
void fo_adv_tra_ver_qr4c(const fo_mesh *m, const double *w, const double *ttf, double num_ord,
                         double *flux, int init_zero);
void fo_adv_tra_ver_cdiff(const fo_mesh *m, const double *w, const double *ttf, double *flux, int init_zero);
void fo_adv_tra_vert_ppm(const fo_mesh *m, double dt, const double *w, const double *ttf, double *flux,
                         int init_zero);
void fo_adv_tra_vert_impl(const fo_mesh *m, double dt, const double *w, double *ttf);
void fo_oce_tra_adv_fct(const fo_mesh *m, double dt, const double *ttf, const double *lo,
                        double *adf_h, double *adf_v, double *ttf_min, double *ttf_max,
                        double *fct_plus, double *fct_minus);
void fo_oce_tra_adv_flux2dtracer(const fo_mesh *m, double dt, double *dttf_h, double *dttf_v,
                                 const double *flux_h, const double *flux_v,
                                 int use_lo, const double *ttf, const double *lo);
int fo_do_oce_adv_tra(const fo_mesh *m, const fo_dyn *d, double dt, fo_tracer *t, fo_work *w);
void fo_tracer_gradient_z(const fo_mesh *m, const double *ttf, double *tr_z);
void fo_diff_part_hor_redi(const fo_mesh *m, const fo_dyn *d, double dt, const double *tr_xy, const double *tr_z,
                           double *del_ttf);
void fo_diff_ver_part_expl_ale(const fo_mesh *m, const fo_dyn *d, double dt, const fo_tracer *t, double *del_ttf);
void fo_diff_ver_part_redi_expl(const fo_mesh *m, const fo_dyn *d, double dt, const double *tr_xy,
                                double *del_ttf);
void fo_diff_tracers_ale_tstar(const fo_mesh *m, const fo_dyn *d, double dt, fo_tracer *t, fo_work *w);
int fo_diff_ver_part_impl_ale(const fo_mesh *m, const fo_dyn *d, double dt, fo_tracer *t);
int fo_bc_surface(const fo_mesh *m, int id, double dt, const fo_dyn *d, int n, int nzmin, double sval, double *out);
int fo_solve_tracers_ale(const fo_mesh *m, const fo_dyn *d, double dt, fo_tracer *t, int ntr, fo_work *w);

int fo_init_stiff_mat_ale(const fo_mesh *m, double dt, int *rowptr, int *colind, double *values, int max_nnz);
int fo_update_stiff_mat_ale(const fo_mesh *m, double dt, const double *dhe, const int *rowptr, const int *colind,
                            double *values);
int fo_ssh_solve_preconditioner(int n, const int *rowptr, const int *colind, const double *values,
                                int variant, double *pr_values);
typedef struct {
    int iters_last, converged, nbreakdown, nnegrz;
    double resid_last, rtol_last;
} fo_cg_stats;
int fo_ssh_solve_cg(int n, const int *rowptr, const int *colind, const double *values,
                    const double *pr_values, double *x, const double *rhs, double soltol,
                    int maxiter, fo_cg_stats *st);

/* vert_vel_ale for which_ale 0 (linfs), 1 (zlevel) or 2 (zstar), with
 * compute_CFLz and compute_Wvel_split; hnode_new is updated in place under
 * zlevel and zstar. cfl_z is in/out: zlevel reads the previous step's CFL_z.
 * zbar (nl) are the reference interface depths (mesh%zbar), used with
 * min_hnode and lzstar_lev by zlevel only. */
int fo_vert_vel_ale(const fo_mesh *m, const double *uv, double dt, int which_ale, const double *hbar,
                    const double *hbar_old, const double *water_flux, int use_wsplit, double wsplit_maxcfl,
                    double *w, double *w_e, double *w_i, double *cfl_z, double *hnode_new, double *cflmax,
                    const double *zbar, double min_hnode, int lzstar_lev);
/* compute_hbar_ale: ssh_rhs_old = column transport divergence (minus the
 * freshwater flux off linfs), hbar_old = hbar, hbar += ssh_rhs_old dt/areasvol,
 * dhe per element; hbar is updated in place */
int fo_compute_hbar_ale(const fo_mesh *m, const double *uv, const double *water_flux, double dt, int which_ale,
                        double *hbar, double *hbar_old, double *ssh_rhs_old, double *dhe);
int fo_compute_ssh_rhs_ale(const fo_mesh *m, const double *uv, const double *uv_rhs, const double *ssh_rhs_old,
                           const double *water_flux, double alpha, int which_ale, int use_cavity_fw2press,
                           double *ssh_rhs);

#ifdef __cplusplus
}
#endif
#endif
