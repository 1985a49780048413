/*
 * fesom_gpu.h -- C-ABI of the MI355X (gfx950) FESOM2 tracer-transport / SSH library.
 *
 * One context = one FESOM2 partition resident on one GPU. The mesh is uploaded
 * once; fields stay in HBM between calls. Arrays are passed exactly as the
 * Fortran derived types hold them: column-major, level index fastest, FP64,
 * 1-based connectivity (edge_tri(2,e) <= 0 marks a boundary edge).
 *
 * Every function returns FGPU_OK (0) or a negative status; fgpu_last_error()
 * gives the message of the last failure in the calling thread. A status of
 * FGPU_ERR_SCHEME / FGPU_ERR_BC corresponds to the reference's par_ex abort
 * (unknown advection scheme, unknown boundary-condition tracer ID); the caller
 * decides whether to abort.
 *
 * The reference interfaces each entry point replaces are named beside it
 * (paths relative to the FESOM2 source tree).
 */
#ifndef FESOM_GPU_H
#define FESOM_GPU_H

#ifdef __cplusplus
extern "C" {
#endif

#define FGPU_OK 0
#define FGPU_ERR_ARG (-1)
#define FGPU_ERR_HIP (-2)
#define FGPU_ERR_SCHEME (-3)
#define FGPU_ERR_BC (-4)
#define FGPU_ERR_STATE (-5)
#define FGPU_ERR_NOMEM (-6)

typedef struct fgpu_context fgpu_context;

typedef struct {
    int iters_last;      /* solverinfo%iters_last   (MOD_DYN.F90:13-46) */
    int converged;       /* 1 if ||r||/||b|| < soltol was reached */
    int nbreakdown;      /* solverinfo%nbreakdown   (solver.F90:259-266, 327-333) */
    int nnegrz;          /* solverinfo%nnegrz       (solver.F90:319-325) */
    double resid_last;   /* sqrt(r.r / nod2D)       (solver.F90:362) */
    double rtol_last;    /* soltol*sqrt(b.b/nod2D)  (solver.F90:180) */
    double ms;           /* device time of the solve (hipEvent), milliseconds */
} fgpu_cg_stats;

const char *fgpu_last_error(void);
int fgpu_version(void);

/* ---- context ------------------------------------------------------------ */
int fgpu_create(fgpu_context **ctx, int device);
int fgpu_destroy(fgpu_context *ctx);
int fgpu_synchronize(fgpu_context *ctx);

/* ---- mesh: replaces the device side of mesh_setup (oce_mesh.F90:153) -----
 * Upload-once 2-D connectivity, geometry and level numbers of t_mesh
 * (MOD_MESH.F90:22-175). coord_nod2D(2,nod2D) radians; elem2D_nodes(3,elem2D),
 * edges(2,edge2D), edge_tri(2,edge2D), nod_in_elem2D(kmax,nod2D) 1-based;
 * nlevels/ulevels(elem2D), nlevels_nod2D/ulevels_nod2D/nlevels_nod2D_min/
 * ulevels_nod2D_max(nod2D); elem_area, elem_cos(elem2D); edge_dxdy(2,edge2D);
 * edge_cross_dxdy(4,edge2D); gradient_sca(6,elem2D).
 * Also runs find_up_downwind_triangles (oce_muscl_adv.F90:162-352) and builds
 * the node->edge gather lists and the SSH sparsity (oce_ale.F90:1440-1496). */
int fgpu_mesh_init(fgpu_context *ctx, int nod2D, int elem2D, int edge2D, int nl, int kmax,
                   double cyclic_length, const double *coord_nod2D, const int *elem2D_nodes,
                   const int *edges, const int *edge_tri, const int *nod_in_elem2D,
                   const int *nod_in_elem2D_num, const int *nlevels, const int *ulevels,
                   const int *nlevels_nod2D, const int *ulevels_nod2D,
                   const int *nlevels_nod2D_min, const int *ulevels_nod2D_max,
                   const double *elem_area, const double *elem_cos, const double *edge_dxdy,
                   const double *edge_cross_dxdy, const double *gradient_sca);

/* 3-D geometry of the current ALE state (linfs: once; zstar: per step).
 * area, areasvol, zbar_3d_n (nl,nod2D); hnode, hnode_new, Z_3d_n (nl-1,nod2D);
 * helem (nl-1,elem2D); zbar_n_bot(nod2D); zbar_e_bot, zbar_e_srf(elem2D). */
int fgpu_set_vertical(fgpu_context *ctx, const double *area, const double *areasvol,
                      const double *hnode, const double *hnode_new, const double *helem,
                      const double *zbar_3d_n, const double *Z_3d_n, const double *zbar_n_bot,
                      const double *zbar_e_bot, const double *zbar_e_srf);

/* ---- dynamics inputs: the vel/w/wi/we arguments of do_oce_adv_tra --------
 * (oce_adv_tra_driver.F90:46-66); uv(2,nl-1,elem2D); w, w_e, w_i (nl,nod2D).
 * use_wsplit = 1 (dynamics%use_wsplit): W = w_e + w_i with w_i taken
 * implicitly, in the LO solution of FCT tracers (adv_tra_vert_impl,
 * oce_adv_tra_ver.F90:90-240, called at oce_adv_tra_driver.F90:323-331) and in
 * the implicit vertical diffusion of the others (oce_ale_tracer.F90:1118-1202). */
int fgpu_set_velocity(fgpu_context *ctx, const double *uv, const double *w, const double *w_e,
                      const double *w_i, int use_wsplit);
/* o_ARRAYS Ki(nl-1,nod2D) and slope_tapered(3,nl-1,nod2D), and Redi of
 * oce_modules.F90:75: the inputs of diff_part_hor_redi
 * (oce_ale_tracer.F90:1526-1689), diff_ver_part_redi_expl (:1439-1522) and the
 * K33 = Kv + Ki S3^2 terms of diff_ver_part_impl_ale (:1105-1184).
 * Ki = NULL means Ki = 0: the horizontal diffusion adds nothing (the default
 * until this is called). slope_tapered may be NULL when redi = 0. */
int fgpu_set_ki(fgpu_context *ctx, const double *Ki, const double *slope_tapered, int redi);
/* o_ARRAYS%Kv(nl,nod2D), read by diff_ver_part_impl_ale (oce_ale_tracer.F90:914) */
int fgpu_set_kv(fgpu_context *ctx, const double *Kv);
/* surface fluxes read by bc_surface (oce_ale_tracer.F90:1873-1881), (nod2D) each */
int fgpu_set_surface_forcing(fgpu_context *ctx, const double *heat_flux, const double *water_flux,
                             const double *virtual_salt, const double *relax_salt,
                             const double *real_salt_flux, double is_nonlinfs);

/* Per-node surface fluxes of the tracer IDs 101-104 and 304 in bc_surface
 * (oce_ale_tracer.F90:2059-2079): prec_rain(nod2D) (g_forcing_arrays, ID 101);
 * wiso_flux_oce(nod2D,3) (o_ARRAYS, node index fastest; IDs 102, 103, 104 take
 * columns 1, 2, 3); hosing_flux(nod2D) and hosing_flux3D(nl-1,nod2D) (o_ARRAYS,
 * ID 304: hosing_depth = use_hosing .and. hosing_mode == 'depth' takes
 * hosing_flux3D at the column's first wet level, else hosing_flux). A NULL
 * array is left as it was; a step whose tracers need an array never set fails
 * with FGPU_ERR_STATE. */
int fgpu_set_surface_forcing_tracers(fgpu_context *ctx, const double *prec_rain, const double *wiso_flux_oce,
                                     const double *hosing_flux, const double *hosing_flux3D, int hosing_depth);

/* ---- the steps either side of the transport -------------------------------
 * which_ALE of g_config (gen_modules_config.F90): */
#define FGPU_ALE_LINFS 0
#define FGPU_ALE_ZLEVEL 1
#define FGPU_ALE_ZSTAR 2

/* which_ALE = 'zlevel' settings (gen_modules_config.F90:69-80) and
 * mesh%zbar (nl), the nominal interface depths: a falling surface that would
 * thin the surface layer below min_hnode of its nominal thickness is spread
 * over the top lzstar_lev layers (1..32). */
int fgpu_set_ale_config(fgpu_context *ctx, const double *zbar, double min_hnode, int lzstar_lev);

/* dynamics%uv (2,nl-1,elem2D) alone, for fgpu_vert_vel_ale and
 * fgpu_compute_ssh_rhs_ale (fgpu_set_velocity also sets it) */
int fgpu_set_uv(fgpu_context *ctx, const double *uv);

/* vert_vel_ale (oce_ale.F90:2107-2668) followed by compute_CFLz
 * (oce_ale.F90:2910-2997) and compute_Wvel_split (oce_ale.F90:3003-3049), on
 * the device from the resident uv: dynamics%w, w_e, w_i and cfl_z stay in HBM
 * for the transport (fgpu_vert_vel_download copies them out). which_ALE linfs,
 * zlevel (oce_ale.F90:2336-2420; needs fgpu_set_ale_config and reads the
 * previous call's CFL_z, zero at first) or zstar; zlevel and zstar also read
 * hbar, hbar_old (nod2D, host; both NULL: the
 * device's from fgpu_compute_hbar_ale) and the water flux
 * of fgpu_set_surface_forcing, and updates hnode_new and the implicit
 * diffusion's Z levels. Levels outside [ulevels_nod2D, nlevels_nod2D] of w_e
 * and w_i are written as zero. use_wsplit = 1 splits W where CFL_z exceeds
 * wsplit_maxcfl and turns on the implicit w_i terms of the transport (as
 * fgpu_set_velocity with use_wsplit = 1). cflmax: max of CFL_z (may be NULL); nan_columns: columns with a
 * NaN in W, which the reference reports (may be NULL). */
int fgpu_vert_vel_ale(fgpu_context *ctx, double dt, int which_ale, const double *hbar, const double *hbar_old,
                      int use_wsplit, double wsplit_maxcfl, double *cflmax, int *nan_columns);
/* w, w_e, w_i, cfl_z (nl,nod2D) and hnode_new (nl-1,nod2D); NULL skips one */
int fgpu_vert_vel_download(fgpu_context *ctx, double *w, double *w_e, double *w_i, double *cfl_z,
                           double *hnode_new);

/* compute_ssh_rhs_ale (oce_ale.F90:1803-1938) from the resident uv and
 * uv_rhs (2,nl-1,elem2D), ssh_rhs_old (nod2D) and alpha; the full free surface
 * (which_ale != FGPU_ALE_LINFS) also reads the water flux of
 * fgpu_set_surface_forcing. The result becomes the SSH solve's right-hand side
 * when the SSH system is set up (fgpu_ssh_solve_resident then solves it) and is
 * copied to ssh_rhs (nod2D) unless that is NULL. ssh_rhs_old = NULL takes the
 * device's, written by the last fgpu_compute_hbar_ale (the reference's
 * dynamics%ssh_rhs_old). */
int fgpu_compute_ssh_rhs_ale(fgpu_context *ctx, const double *uv_rhs, const double *ssh_rhs_old, double alpha,
                             int which_ale, int use_cavity_fw2press, double *ssh_rhs);

/* compute_hbar_ale (oce_ale.F90:1956-2091) from the resident uv:
 * ssh_rhs_old = the column transport divergence (minus water_flux * areasvol
 * at open-ocean nodes off linfs), hbar_old = hbar, hbar += ssh_rhs_old * dt /
 * areasvol at open-ocean nodes, dhe per element (the mean hbar change, 0 under
 * cavities). hbar (nod2D) may be NULL to continue from the device's hbar of
 * the previous call. The four results stay on the device for
 * fgpu_vert_vel_ale (zstar, hbar and hbar_old NULL), fgpu_update_stiff_mat_ale
 * (dhe NULL) and fgpu_compute_ssh_rhs_ale (ssh_rhs_old NULL); each *_out that
 * is not NULL receives a copy (hbar, hbar_old, ssh_rhs_old: nod2D; dhe: elem2D). */
int fgpu_compute_hbar_ale(fgpu_context *ctx, double dt, int which_ale, const double *hbar, double *hbar_out,
                          double *hbar_old_out, double *ssh_rhs_old_out, double *dhe_out);

/* ---- tracers: t_tracer (MOD_TRACER.F90:10-109) ---------------------------
 * num_tracers tracers; each configured with its namelist entry
 * (config/namelist.tra:33-38): hor in {"MFCT","UPW1","MUSCL"}, ver in {"QR4C","UPW1","CDIFF","PPM"},
 * lim in {"FCT","NON"} (trailing blanks ignored), tra_adv_ph/pv, AB_order 2|3,
 * i_vert_diff. tr_num is 1-based like the reference. ID selects bc_surface
 * (oce_ale_tracer.F90:1862-2098): 1 temperature, 2 salinity, 100 and 301-303
 * passive (no surface flux), 101 (prec_rain), 102-104 (wiso_flux_oce) and 304
 * (hosing) with the fluxes of fgpu_set_surface_forcing_tracers, 501 ice-shelf
 * water (dt*water_flux*(T-1) under a cavity); any other ID fails with
 * FGPU_ERR_BC, as the reference stops. */
int fgpu_tracers_init(fgpu_context *ctx, int num_tracers);
int fgpu_tracer_config(fgpu_context *ctx, int tr_num, int ID, int AB_order, const char *tra_adv_hor,
                       const char *tra_adv_ver, const char *tra_adv_lim, double tra_adv_ph,
                       double tra_adv_pv, int i_vert_diff);
/* values(nl-1,nod2D); valuesold(AB_order-1,nl-1,nod2D) */
int fgpu_tracer_upload(fgpu_context *ctx, int tr_num, const double *values, const double *valuesold);
int fgpu_tracer_download(fgpu_context *ctx, int tr_num, double *values, double *valuesold);
/* Download a work array of the last per-tracer call: name is one of
 * "valuesAB", "tr_xy", "edge_up_dn_grad", "fct_LO", "adv_flux_hor",
 * "adv_flux_ver", "del_ttf_advhoriz", "del_ttf_advvert", "fct_plus", "fct_minus".
 * Shapes follow t_tracer_work (MOD_TRACER.F90:33-67). */
int fgpu_tracer_work_download(fgpu_context *ctx, int tr_num, const char *name, double *out, long n);
/* edge_up_dn_tri(2,edge2D), 1-based, 0 = none */
int fgpu_edge_up_dn_tri(fgpu_context *ctx, int *out);

/* valuesAB(nl-1,nod2D) of tracer tr_num as the caller's init_tracers_AB left
 * it (oce_tracer_mod.F90:44-74); the element gradients of the resident values
 * (tracer_gradient_elements, oce_tracer_mod.F90:126-131) are recomputed, so
 * fgpu_do_oce_adv_tra then sees what the reference's do_oce_adv_tra sees
 * after init_tracers_AB on the host. */
int fgpu_tracer_set_valuesAB(fgpu_context *ctx, int tr_num, const double *valuesAB);

/* ---- per-tracer entry points (reference call surface) --------------------- */
/* init_tracers_AB (oce_tracer_mod.F90:13-143) */
int fgpu_init_tracers_AB(fgpu_context *ctx, int tr_num);
/* do_oce_adv_tra (oce_adv_tra_driver.F90:46-490): leaves del_ttf_advhoriz and
 * del_ttf_advvert in device memory, like tracers%work. */
int fgpu_do_oce_adv_tra(fgpu_context *ctx, double dt, int tr_num);
/* diff_tracers_ale (oce_ale_tracer.F90:533-793) incl. the del_ttf sum of
 * solve_tracers_ale (:318-322): T* update and, if i_vert_diff, implicit diffusion */
int fgpu_diff_tracers_ale(fgpu_context *ctx, double dt, int tr_num);
/* diff_ver_part_impl_ale (oce_ale_tracer.F90:907-1435) alone */
int fgpu_diff_ver_part_impl_ale(fgpu_context *ctx, double dt, int tr_num);

/* ---- batched step: replaces the tracer loop of solve_tracers_ale
 * (oce_ale_tracer.F90:280-385) and the salinity clamp (:502-514). All tracers
 * advance together; nsteps > 1 replays the captured step graph. */
int fgpu_solve_tracers_ale(fgpu_context *ctx, double dt, int nsteps);
/* Same, timed with hipEvents on the library stream; *ms = elapsed device time. */
int fgpu_solve_tracers_ale_timed(fgpu_context *ctx, double dt, int nsteps, double *ms);
/* Per-kernel timing of the batched step: when enabled, every kernel launch of
 * fgpu_solve_tracers_ale is bracketed by hipEvents on the library stream. */
int fgpu_profile_enable(fgpu_context *ctx, int on);
int fgpu_profile_count(fgpu_context *ctx);
int fgpu_profile_read(fgpu_context *ctx, int idx, char *name, int name_len, double *total_ms, int *launches);
/* Autotuner of the fused step (no counterpart in the reference): times the
 * candidate launch shapes (tracers per block 1, 2, 4, ..., all, crossed with 2,
 * 4, 5 or 8 tracers per Thomas chunk and 1, 4 or 8 columns per block of the last
 * kernel for columns of up to 64 levels, 1 to 4 for columns of up to 72
 * interfaces, 1 or 2 for deeper ones) over nsteps
 * replayed steps, keeps the fastest
 * for this context and reports its tracers per block and time. The tracer
 * state is restored afterwards; results do not depend on the choice. */
int fgpu_autotune_step(fgpu_context *ctx, double dt, int nsteps, int *best_tpb, double *best_ms_per_step);
/* The fused step's current launch shape (no counterpart in the reference):
 * tracers per block, tracers per Thomas chunk, columns per block of the last
 * kernel as launched (a shape that needs another chunk size reports the one used). */
int fgpu_step_shape(fgpu_context *ctx, int *tracers_per_block, int *thomas_chunk, int *columns_per_block);
/* Kernel variants of the step (no counterpart in the reference). kd_variant,
 * the last kernel of the plain step:
 * 0 the block's waves in lock-step, the sweeps on the first wave's lanes;
 * 1 one wave per column plus a wave per block that sweeps the previous chunk
 *   while the column waves compute the next;
 * 2 one wave per column that sweeps its own column.
 * kc_tracers, the bound kernel: 0 one column's levels over the lanes, 1, 2 or 4
 * tracers per block with (tracer, level) items over the lanes.
 * ka_levels_per_lane, the edge-flux kernel: 1 one level per lane (64 or 128
 * lanes per edge), 2 two adjacent levels per lane with 16-byte accesses (one
 * wave per edge; meshes with an even number of layers only).
 * kb_levels_per_lane, the LO kernel, likewise (also not with the W split or
 * CDIFF/PPM tracers).
 * The autotuner picks all four with the shape; FGPU_STEP_KD=lockstep|sweepwave|
 * wavelocal, FGPU_STEP_KC=0|1|2|4, FGPU_STEP_KA=1|2 and FGPU_STEP_KB=1|2 set
 * them when a context is created. Results are the same bit for bit. */
int fgpu_step_variant(fgpu_context *ctx, int *kd_variant, int *kc_tracers, int *ka_levels_per_lane,
                      int *kb_levels_per_lane);
/* Sets the launch shape and the kernel variants instead of the autotuner (no
 * counterpart in the reference); values outside the sets above fall back as
 * fgpu_step_shape reports. */
int fgpu_set_step_shape(fgpu_context *ctx, int tracers_per_block, int thomas_chunk, int columns_per_block,
                        int kd_variant, int kc_tracers, int ka_levels_per_lane, int kb_levels_per_lane);
/* Tracer content sum_n sum_k areasvol*hnode_new*T (deterministic device sum). */
int fgpu_tracer_content(fgpu_context *ctx, int tr_num, double *content);

/* ---- SSH solve ------------------------------------------------------------ */
/* init_stiff_mat_ale (oce_ale.F90:1393-1582): assemble on the device from the
 * uploaded mesh (factor g*dt*alpha*theta). */
int fgpu_init_stiff_mat_ale(fgpu_context *ctx, double dt, double g, double alpha, double theta);

/* update_stiff_mat_ale (oce_ale.F90:1683-1792), full free surface: adds the
 * stiffness term of the ssh change dhe (elem2D) to the matrix assembled by
 * fgpu_init_stiff_mat_ale, with factor g*dt*alpha*theta. Like the reference,
 * the preconditioner is not rebuilt (call fgpu_ssh_solve_preconditioner to
 * refresh it). dhe = NULL takes the device's, from fgpu_compute_hbar_ale. */
int fgpu_update_stiff_mat_ale(fgpu_context *ctx, const double *dhe, double dt, double g, double alpha, double theta);

/* Alternatively hand over an assembled ssh_stiff (MOD_MESH.F90:11-20), 1-based,
 * diagonal first in every row. */
int fgpu_ssh_set_matrix(fgpu_context *ctx, int dim, int nza, const int *rowptr, const int *colind,
                        const double *values);
int fgpu_ssh_get_matrix(fgpu_context *ctx, int *rowptr, int *colind, double *values,
                        double *pr_values, int *nza);
/* ssh_solve_preconditioner (solver.F90:31-117): variant 0 and 1 as in the
 * reference, 2 = diagonal (Jacobi) only. */
int fgpu_ssh_solve_preconditioner(fgpu_context *ctx, int precond_variant);
/* mesh%ssh_stiff%pr_values(nza) as built by the caller's ssh_solve_preconditioner
 * (solver.F90:31-117, same sparsity as the matrix) with its precond_variant;
 * replaces the device build of fgpu_ssh_solve_preconditioner. */
int fgpu_ssh_set_preconditioner(fgpu_context *ctx, int precond_variant, const double *pr_values);
/* ssh_solve_cg (solver.F90:120-385): x (nod2D) in/out (warm start), rhs (nod2D). */
int fgpu_ssh_solve_cg(fgpu_context *ctx, double *x, const double *rhs, double soltol, int maxiter,
                      fgpu_cg_stats *stats);
/* Device-resident variant for repeated solves: upload rhs/x0 once with
 * fgpu_ssh_upload, solve nrep times from the same x0, download with fgpu_ssh_download. */
int fgpu_ssh_upload(fgpu_context *ctx, const double *x0, const double *rhs);
int fgpu_ssh_solve_resident(fgpu_context *ctx, double soltol, int maxiter, int nrep,
                            fgpu_cg_stats *stats);
int fgpu_ssh_download(fgpu_context *ctx, double *x);
/* No counterpart in the reference: fgpu_ssh_solve_cg for callers that cannot
 * allocate device memory (code under graph capture, tight serving loops,
 * harnesses that forbid allocation). x_dev (in/out) and rhs_dev are the
 * caller's DEVICE vectors of nod2D entries on the context's device; the work
 * vectors come from a static device area of FGPU_NOALLOC_MAX_ROWS rows (about
 * 108 MB on each GPU the library runs on, held until the process ends), so the
 * call allocates nothing. Larger systems return FGPU_ERR_ARG. Same kernels and
 * iteration order as fgpu_ssh_solve_cg: the same x bit for bit. Calls that
 * overlap on one GPU share the static area and race. stats->ms is host time. */
#define FGPU_NOALLOC_MAX_ROWS 1500000
int fgpu_ssh_solve_cg_noalloc(fgpu_context *ctx, double *x_dev, const double *rhs_dev, double soltol, int maxiter,
                              fgpu_cg_stats *stats);

/* ---- learned SSH preconditioner -------------------------------------------
 * No counterpart in the reference: an optional replacement of the diagonal /
 * sparse M^-1 of ssh_solve_cg (solver.F90:120-385) by a message-passing network
 * on the SSH matrix graph (3 layers, width 64, bf16 weights and activations
 * with FP32 accumulation) inside flexible CG (FCG(1)). Vectors, residual and
 * dot products stay FP64; the stopping test is the reference's
 * sqrt(r.r/nod2D) < soltol*sqrt(b.b/nod2D). A solve that breaks down or reaches
 * maxiter continues with plain diagonally preconditioned CG from its x for up to
 * fallback_maxiter iterations. All calls need the SSH matrix (and, for the
 * solves, fgpu_ssh_solve_preconditioner) first.
 * Parameters: FP32 vector of fgpu_ssh_nn_param_count() entries, in the order
 * W0(2,64) | W_l(128,64) b_l(64), l = 1..3 | w_out(64) b_out(1), row-major. */
int fgpu_ssh_nn_param_count(void);
/* He-initialised hidden layers, zero output layer (= the diagonal preconditioner) */
int fgpu_ssh_nn_init(fgpu_context *ctx, unsigned long long seed);
int fgpu_ssh_nn_set_weights(fgpu_context *ctx, const float *theta, int np);
int fgpu_ssh_nn_get_weights(fgpu_context *ctx, float *theta, int np);
/* z = M(r), nod2D entries each */
int fgpu_ssh_nn_apply(fgpu_context *ctx, const double *r, double *z);
/* Training loss mean_s (M(b_s) - x_s)^T A (M(b_s) - x_s) / (x_s^T A x_s) over nsamp
 * right-hand sides b (nsamp x nod2D) with solutions x, and its gradient (FP32, np entries). */
int fgpu_ssh_nn_loss_grad(fgpu_context *ctx, const double *b, const double *x, int nsamp, double *loss,
                          float *grad);
/* Adam (lr, beta 0.9/0.999) over mini-batches of `batch` samples for `epochs`
 * passes; loss_per_epoch[epochs] receives the mean loss of each pass. */
int fgpu_ssh_nn_train(fgpu_context *ctx, const double *b, const double *x, int nsamp, int epochs, int batch,
                      double lr, double *loss_per_epoch);
/* learned FCG solve; stats->iters_last counts FCG plus fallback iterations,
 * *fallback_iters the fallback's share; stats->nnegrz counts the iterations in
 * which the network's z had z.r <= 0 and the diagonal preconditioner's z was
 * used instead */
int fgpu_ssh_solve_learned(fgpu_context *ctx, double *x, const double *rhs, double soltol, int maxiter,
                           int fallback_maxiter, fgpu_cg_stats *stats, int *fallback_iters);
/* device-resident repeated solves from the x0/rhs of fgpu_ssh_upload (timing) */
int fgpu_ssh_solve_learned_resident(fgpu_context *ctx, double soltol, int maxiter, int fallback_maxiter, int nrep,
                                    fgpu_cg_stats *stats, int *fallback_iters);

#ifdef __cplusplus
}
#endif
#endif
