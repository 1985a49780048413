! fesom_gpu_c.f90 -- ISO_C_BINDING interfaces of libfesom_gpu.so (include/fesom_gpu.h).
!
! One interface per C entry point, same names, same argument order. Arrays are
! passed by address (Fortran default), scalars by value, C strings as
! NUL-terminated character(kind=c_char) arrays; the context is an opaque c_ptr.
! The reference routine each entry point replaces is named in the C header.
module fesom_gpu_c
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: FGPU_OK = 0, FGPU_ERR_ARG = -1, FGPU_ERR_HIP = -2, &
       FGPU_ERR_SCHEME = -3, FGPU_ERR_BC = -4, FGPU_ERR_STATE = -5, FGPU_ERR_NOMEM = -6
  ! which_ALE (g_config)
  integer(c_int), parameter, public :: FGPU_ALE_LINFS = 0, FGPU_ALE_ZLEVEL = 1, FGPU_ALE_ZSTAR = 2

  ! solverinfo counters of one CG solve (MOD_DYN.F90:13-46)
  type, bind(C), public :: fgpu_cg_stats
     integer(c_int) :: iters_last
     integer(c_int) :: converged
     integer(c_int) :: nbreakdown
     integer(c_int) :: nnegrz
     real(c_double) :: resid_last
     real(c_double) :: rtol_last
     real(c_double) :: ms
  end type fgpu_cg_stats

  public :: fgpu_last_error, fgpu_version, fgpu_create, fgpu_destroy, fgpu_synchronize, &
       fgpu_mesh_init, fgpu_set_vertical, fgpu_set_velocity, fgpu_set_kv, fgpu_set_ki, fgpu_set_surface_forcing, &
       fgpu_tracers_init, fgpu_tracer_config, fgpu_tracer_upload, fgpu_tracer_download, &
       fgpu_tracer_work_download, fgpu_edge_up_dn_tri, fgpu_init_tracers_AB, fgpu_do_oce_adv_tra, &
       fgpu_diff_tracers_ale, fgpu_diff_ver_part_impl_ale, fgpu_solve_tracers_ale, &
       fgpu_solve_tracers_ale_timed, fgpu_profile_enable, fgpu_profile_count, fgpu_profile_read, &
       fgpu_tracer_content, fgpu_init_stiff_mat_ale, fgpu_ssh_set_matrix, fgpu_ssh_get_matrix, &
       fgpu_ssh_solve_preconditioner, fgpu_ssh_solve_cg, fgpu_ssh_upload, fgpu_ssh_solve_resident, &
       fgpu_ssh_download, fgpu_autotune_step, fgpu_step_shape, fgpu_ssh_nn_param_count, fgpu_ssh_nn_init, &
       fgpu_ssh_nn_set_weights, fgpu_ssh_nn_get_weights, fgpu_ssh_nn_apply, fgpu_ssh_nn_loss_grad, &
       fgpu_ssh_nn_train, fgpu_ssh_solve_learned, fgpu_ssh_solve_learned_resident, fgpu_set_uv, &
       fgpu_vert_vel_ale, fgpu_vert_vel_download, fgpu_compute_ssh_rhs_ale, fgpu_update_stiff_mat_ale, &
       fgpu_compute_hbar_ale, fgpu_set_ale_config, fgpu_set_surface_forcing_tracers, &
       fgpu_tracer_set_valuesAB, fgpu_ssh_set_preconditioner, fgpu_ssh_solve_cg_noalloc, fgpu_step_variant, &
       fgpu_set_step_shape

  interface
     type(c_ptr) function fgpu_last_error() bind(C, name="fgpu_last_error")
       import :: c_ptr
     end function fgpu_last_error

     integer(c_int) function fgpu_version() bind(C, name="fgpu_version")
       import :: c_int
     end function fgpu_version

     integer(c_int) function fgpu_create(ctx, device) bind(C, name="fgpu_create")
       import :: c_int, c_ptr
       type(c_ptr), intent(out) :: ctx
       integer(c_int), value :: device
     end function fgpu_create

     integer(c_int) function fgpu_destroy(ctx) bind(C, name="fgpu_destroy")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fgpu_destroy

     integer(c_int) function fgpu_synchronize(ctx) bind(C, name="fgpu_synchronize")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fgpu_synchronize

     integer(c_int) function fgpu_mesh_init(ctx, nod2D, elem2D, edge2D, nl, kmax, cyclic_length, &
          coord_nod2D, elem2D_nodes, edges, edge_tri, nod_in_elem2D, nod_in_elem2D_num, nlevels, &
          ulevels, nlevels_nod2D, ulevels_nod2D, nlevels_nod2D_min, ulevels_nod2D_max, elem_area, &
          elem_cos, edge_dxdy, edge_cross_dxdy, gradient_sca) bind(C, name="fgpu_mesh_init")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: nod2D, elem2D, edge2D, nl, kmax
       real(c_double), value :: cyclic_length
       real(c_double), intent(in) :: coord_nod2D(*)
       integer(c_int), intent(in) :: elem2D_nodes(*), edges(*), edge_tri(*), nod_in_elem2D(*), &
            nod_in_elem2D_num(*), nlevels(*), ulevels(*), nlevels_nod2D(*), ulevels_nod2D(*), &
            nlevels_nod2D_min(*), ulevels_nod2D_max(*)
       real(c_double), intent(in) :: elem_area(*), elem_cos(*), edge_dxdy(*), edge_cross_dxdy(*), &
            gradient_sca(*)
     end function fgpu_mesh_init

     integer(c_int) function fgpu_set_vertical(ctx, area, areasvol, hnode, hnode_new, helem, &
          zbar_3d_n, Z_3d_n, zbar_n_bot, zbar_e_bot, zbar_e_srf) bind(C, name="fgpu_set_vertical")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: area(*), areasvol(*), hnode(*), hnode_new(*), helem(*), &
            zbar_3d_n(*), Z_3d_n(*), zbar_n_bot(*), zbar_e_bot(*), zbar_e_srf(*)
     end function fgpu_set_vertical

     integer(c_int) function fgpu_set_velocity(ctx, uv, w, w_e, w_i, use_wsplit) &
          bind(C, name="fgpu_set_velocity")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: uv(*), w(*), w_e(*), w_i(*)
       integer(c_int), value :: use_wsplit
     end function fgpu_set_velocity

     integer(c_int) function fgpu_set_kv(ctx, Kv) bind(C, name="fgpu_set_kv")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: Kv(*)
     end function fgpu_set_kv

     ! Ki = c_null_ptr: Ki = 0 (no horizontal diffusion)
     integer(c_int) function fgpu_set_ki(ctx, Ki, slope_tapered, redi) bind(C, name="fgpu_set_ki")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, Ki, slope_tapered
       integer(c_int), value :: redi
     end function fgpu_set_ki

     integer(c_int) function fgpu_set_surface_forcing(ctx, heat_flux, water_flux, virtual_salt, &
          relax_salt, real_salt_flux, is_nonlinfs) bind(C, name="fgpu_set_surface_forcing")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: heat_flux(*), water_flux(*), virtual_salt(*), relax_salt(*), &
            real_salt_flux(*)
       real(c_double), value :: is_nonlinfs
     end function fgpu_set_surface_forcing

     ! arrays by address: c_null_ptr leaves one unset
     integer(c_int) function fgpu_set_surface_forcing_tracers(ctx, prec_rain, wiso_flux_oce, hosing_flux, &
          hosing_flux3D, hosing_depth) bind(C, name="fgpu_set_surface_forcing_tracers")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx, prec_rain, wiso_flux_oce, hosing_flux, hosing_flux3D
       integer(c_int), value :: hosing_depth
     end function fgpu_set_surface_forcing_tracers

     integer(c_int) function fgpu_tracer_set_valuesAB(ctx, tr_num, valuesAB) bind(C, name="fgpu_tracer_set_valuesAB")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: tr_num
       real(c_double), intent(in) :: valuesAB(*)
     end function fgpu_tracer_set_valuesAB

     integer(c_int) function fgpu_ssh_set_preconditioner(ctx, precond_variant, pr_values) &
          bind(C, name="fgpu_ssh_set_preconditioner")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: precond_variant
       real(c_double), intent(in) :: pr_values(*)
     end function fgpu_ssh_set_preconditioner

     integer(c_int) function fgpu_set_uv(ctx, uv) bind(C, name="fgpu_set_uv")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: uv(*)
     end function fgpu_set_uv

     integer(c_int) function fgpu_vert_vel_ale(ctx, dt, which_ale, hbar, hbar_old, use_wsplit, wsplit_maxcfl, &
          cflmax, nan_columns) bind(C, name="fgpu_vert_vel_ale")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt, wsplit_maxcfl
       integer(c_int), value :: which_ale, use_wsplit
       real(c_double), intent(in) :: hbar(*), hbar_old(*)
       real(c_double), intent(out) :: cflmax
       integer(c_int), intent(out) :: nan_columns
     end function fgpu_vert_vel_ale

     integer(c_int) function fgpu_vert_vel_download(ctx, w, w_e, w_i, cfl_z, hnode_new) &
          bind(C, name="fgpu_vert_vel_download")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: w(*), w_e(*), w_i(*), cfl_z(*), hnode_new(*)
     end function fgpu_vert_vel_download

     integer(c_int) function fgpu_compute_ssh_rhs_ale(ctx, uv_rhs, ssh_rhs_old, alpha, which_ale, &
          use_cavity_fw2press, ssh_rhs) bind(C, name="fgpu_compute_ssh_rhs_ale")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: uv_rhs(*), ssh_rhs_old(*)
       real(c_double), value :: alpha
       integer(c_int), value :: which_ale, use_cavity_fw2press
       real(c_double), intent(out) :: ssh_rhs(*)
     end function fgpu_compute_ssh_rhs_ale

     integer(c_int) function fgpu_update_stiff_mat_ale(ctx, dhe, dt, g, alpha, theta) &
          bind(C, name="fgpu_update_stiff_mat_ale")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: dhe(*)
       real(c_double), value :: dt, g, alpha, theta
     end function fgpu_update_stiff_mat_ale

     integer(c_int) function fgpu_set_ale_config(ctx, zbar, min_hnode, lzstar_lev) &
          bind(C, name="fgpu_set_ale_config")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: zbar(*)
       real(c_double), value :: min_hnode
       integer(c_int), value :: lzstar_lev
     end function fgpu_set_ale_config

     integer(c_int) function fgpu_compute_hbar_ale(ctx, dt, which_ale, hbar, hbar_out, hbar_old_out, &
          ssh_rhs_old_out, dhe_out) bind(C, name="fgpu_compute_hbar_ale")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
       integer(c_int), value :: which_ale
       real(c_double), intent(in) :: hbar(*)
       real(c_double), intent(out) :: hbar_out(*), hbar_old_out(*), ssh_rhs_old_out(*), dhe_out(*)
     end function fgpu_compute_hbar_ale

     integer(c_int) function fgpu_tracers_init(ctx, num_tracers) bind(C, name="fgpu_tracers_init")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: num_tracers
     end function fgpu_tracers_init

     integer(c_int) function fgpu_tracer_config(ctx, tr_num, ID, AB_order, tra_adv_hor, tra_adv_ver, &
          tra_adv_lim, tra_adv_ph, tra_adv_pv, i_vert_diff) bind(C, name="fgpu_tracer_config")
       import :: c_int, c_ptr, c_double, c_char
       type(c_ptr), value :: ctx
       integer(c_int), value :: tr_num, ID, AB_order
       character(kind=c_char), intent(in) :: tra_adv_hor(*), tra_adv_ver(*), tra_adv_lim(*)
       real(c_double), value :: tra_adv_ph, tra_adv_pv
       integer(c_int), value :: i_vert_diff
     end function fgpu_tracer_config

     integer(c_int) function fgpu_tracer_upload(ctx, tr_num, values, valuesold) &
          bind(C, name="fgpu_tracer_upload")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: tr_num
       real(c_double), intent(in) :: values(*), valuesold(*)
     end function fgpu_tracer_upload

     integer(c_int) function fgpu_tracer_download(ctx, tr_num, values, valuesold) &
          bind(C, name="fgpu_tracer_download")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: tr_num
       real(c_double), intent(out) :: values(*), valuesold(*)
     end function fgpu_tracer_download

     integer(c_int) function fgpu_tracer_work_download(ctx, tr_num, name, out, n) &
          bind(C, name="fgpu_tracer_work_download")
       import :: c_int, c_ptr, c_double, c_char, c_long
       type(c_ptr), value :: ctx
       integer(c_int), value :: tr_num
       character(kind=c_char), intent(in) :: name(*)
       real(c_double), intent(out) :: out(*)
       integer(c_long), value :: n
     end function fgpu_tracer_work_download

     integer(c_int) function fgpu_edge_up_dn_tri(ctx, out) bind(C, name="fgpu_edge_up_dn_tri")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: out(*)
     end function fgpu_edge_up_dn_tri

     integer(c_int) function fgpu_init_tracers_AB(ctx, tr_num) bind(C, name="fgpu_init_tracers_AB")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: tr_num
     end function fgpu_init_tracers_AB

     integer(c_int) function fgpu_do_oce_adv_tra(ctx, dt, tr_num) bind(C, name="fgpu_do_oce_adv_tra")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
       integer(c_int), value :: tr_num
     end function fgpu_do_oce_adv_tra

     integer(c_int) function fgpu_diff_tracers_ale(ctx, dt, tr_num) bind(C, name="fgpu_diff_tracers_ale")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
       integer(c_int), value :: tr_num
     end function fgpu_diff_tracers_ale

     integer(c_int) function fgpu_diff_ver_part_impl_ale(ctx, dt, tr_num) &
          bind(C, name="fgpu_diff_ver_part_impl_ale")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
       integer(c_int), value :: tr_num
     end function fgpu_diff_ver_part_impl_ale

     integer(c_int) function fgpu_solve_tracers_ale(ctx, dt, nsteps) bind(C, name="fgpu_solve_tracers_ale")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
       integer(c_int), value :: nsteps
     end function fgpu_solve_tracers_ale

     integer(c_int) function fgpu_solve_tracers_ale_timed(ctx, dt, nsteps, ms) &
          bind(C, name="fgpu_solve_tracers_ale_timed")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
       integer(c_int), value :: nsteps
       real(c_double), intent(out) :: ms
     end function fgpu_solve_tracers_ale_timed

     integer(c_int) function fgpu_profile_enable(ctx, on) bind(C, name="fgpu_profile_enable")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: on
     end function fgpu_profile_enable

     integer(c_int) function fgpu_profile_count(ctx) bind(C, name="fgpu_profile_count")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
     end function fgpu_profile_count

     integer(c_int) function fgpu_profile_read(ctx, idx, name, name_len, total_ms, launches) &
          bind(C, name="fgpu_profile_read")
       import :: c_int, c_ptr, c_double, c_char
       type(c_ptr), value :: ctx
       integer(c_int), value :: idx, name_len
       character(kind=c_char), intent(out) :: name(*)
       real(c_double), intent(out) :: total_ms
       integer(c_int), intent(out) :: launches
     end function fgpu_profile_read

     integer(c_int) function fgpu_tracer_content(ctx, tr_num, content) bind(C, name="fgpu_tracer_content")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: tr_num
       real(c_double), intent(out) :: content
     end function fgpu_tracer_content

     integer(c_int) function fgpu_init_stiff_mat_ale(ctx, dt, g, alpha, theta) &
          bind(C, name="fgpu_init_stiff_mat_ale")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt, g, alpha, theta
     end function fgpu_init_stiff_mat_ale

     integer(c_int) function fgpu_ssh_set_matrix(ctx, dim, nza, rowptr, colind, values) &
          bind(C, name="fgpu_ssh_set_matrix")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       integer(c_int), value :: dim, nza
       integer(c_int), intent(in) :: rowptr(*), colind(*)
       real(c_double), intent(in) :: values(*)
     end function fgpu_ssh_set_matrix

     integer(c_int) function fgpu_ssh_get_matrix(ctx, rowptr, colind, values, pr_values, nza) &
          bind(C, name="fgpu_ssh_get_matrix")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       type(c_ptr), value :: rowptr, colind, values, pr_values
       integer(c_int), intent(out) :: nza
     end function fgpu_ssh_get_matrix

     integer(c_int) function fgpu_ssh_solve_preconditioner(ctx, precond_variant) &
          bind(C, name="fgpu_ssh_solve_preconditioner")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: precond_variant
     end function fgpu_ssh_solve_preconditioner

     integer(c_int) function fgpu_ssh_solve_cg(ctx, x, rhs, soltol, maxiter, stats) &
          bind(C, name="fgpu_ssh_solve_cg")
       import :: c_int, c_ptr, c_double, fgpu_cg_stats
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: x(*)
       real(c_double), intent(in) :: rhs(*)
       real(c_double), value :: soltol
       integer(c_int), value :: maxiter
       type(fgpu_cg_stats), intent(out) :: stats
     end function fgpu_ssh_solve_cg

     integer(c_int) function fgpu_ssh_upload(ctx, x0, rhs) bind(C, name="fgpu_ssh_upload")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: x0(*), rhs(*)
     end function fgpu_ssh_upload

     integer(c_int) function fgpu_ssh_solve_resident(ctx, soltol, maxiter, nrep, stats) &
          bind(C, name="fgpu_ssh_solve_resident")
       import :: c_int, c_ptr, c_double, fgpu_cg_stats
       type(c_ptr), value :: ctx
       real(c_double), value :: soltol
       integer(c_int), value :: maxiter, nrep
       type(fgpu_cg_stats), intent(out) :: stats
     end function fgpu_ssh_solve_resident

     integer(c_int) function fgpu_ssh_download(ctx, x) bind(C, name="fgpu_ssh_download")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(out) :: x(*)
     end function fgpu_ssh_download

     ! x_dev, rhs_dev: device addresses (c_ptr) of nod2D doubles; work vectors
     ! from the library's static device area (no allocation)
     integer(c_int) function fgpu_ssh_solve_cg_noalloc(ctx, x_dev, rhs_dev, soltol, maxiter, stats) &
          bind(C, name="fgpu_ssh_solve_cg_noalloc")
       import :: c_int, c_ptr, c_double, fgpu_cg_stats
       type(c_ptr), value :: ctx, x_dev, rhs_dev
       real(c_double), value :: soltol
       integer(c_int), value :: maxiter
       type(fgpu_cg_stats), intent(out) :: stats
     end function fgpu_ssh_solve_cg_noalloc

     integer(c_int) function fgpu_set_step_shape(ctx, tracers_per_block, thomas_chunk, columns_per_block, &
          kd_variant, kc_tracers, ka_levels_per_lane, kb_levels_per_lane) bind(C, name="fgpu_set_step_shape")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), value :: tracers_per_block, thomas_chunk, columns_per_block, kd_variant, kc_tracers
       integer(c_int), value :: ka_levels_per_lane, kb_levels_per_lane
     end function fgpu_set_step_shape

     integer(c_int) function fgpu_step_variant(ctx, kd_variant, kc_tracers, ka_levels_per_lane, &
          kb_levels_per_lane) bind(C, name="fgpu_step_variant")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: kd_variant, kc_tracers, ka_levels_per_lane, kb_levels_per_lane
     end function fgpu_step_variant

     integer(c_int) function fgpu_autotune_step(ctx, dt, nsteps, best_tpb, best_ms) &
          bind(C, name="fgpu_autotune_step")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), value :: dt
       integer(c_int), value :: nsteps
       integer(c_int), intent(out) :: best_tpb
       real(c_double), intent(out) :: best_ms
     end function fgpu_autotune_step

     integer(c_int) function fgpu_step_shape(ctx, tracers_per_block, thomas_chunk, columns_per_block) &
          bind(C, name="fgpu_step_shape")
       import :: c_int, c_ptr
       type(c_ptr), value :: ctx
       integer(c_int), intent(out) :: tracers_per_block, thomas_chunk, columns_per_block
     end function fgpu_step_shape

     integer(c_int) function fgpu_ssh_nn_param_count() bind(C, name="fgpu_ssh_nn_param_count")
       import :: c_int
     end function fgpu_ssh_nn_param_count

     integer(c_int) function fgpu_ssh_nn_init(ctx, seed) bind(C, name="fgpu_ssh_nn_init")
       import :: c_int, c_ptr, c_long_long
       type(c_ptr), value :: ctx
       integer(c_long_long), value :: seed
     end function fgpu_ssh_nn_init

     integer(c_int) function fgpu_ssh_nn_set_weights(ctx, theta, np) bind(C, name="fgpu_ssh_nn_set_weights")
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(in) :: theta(*)
       integer(c_int), value :: np
     end function fgpu_ssh_nn_set_weights

     integer(c_int) function fgpu_ssh_nn_get_weights(ctx, theta, np) bind(C, name="fgpu_ssh_nn_get_weights")
       import :: c_int, c_ptr, c_float
       type(c_ptr), value :: ctx
       real(c_float), intent(out) :: theta(*)
       integer(c_int), value :: np
     end function fgpu_ssh_nn_get_weights

     integer(c_int) function fgpu_ssh_nn_apply(ctx, r, z) bind(C, name="fgpu_ssh_nn_apply")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: r(*)
       real(c_double), intent(out) :: z(*)
     end function fgpu_ssh_nn_apply

     integer(c_int) function fgpu_ssh_nn_loss_grad(ctx, b, x, nsamp, loss, grad) bind(C, name="fgpu_ssh_nn_loss_grad")
       import :: c_int, c_ptr, c_double, c_float
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: b(*), x(*)
       integer(c_int), value :: nsamp
       real(c_double), intent(out) :: loss
       real(c_float), intent(out) :: grad(*)
     end function fgpu_ssh_nn_loss_grad

     integer(c_int) function fgpu_ssh_nn_train(ctx, b, x, nsamp, epochs, batch, lr, loss_per_epoch) &
          bind(C, name="fgpu_ssh_nn_train")
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: ctx
       real(c_double), intent(in) :: b(*), x(*)
       integer(c_int), value :: nsamp, epochs, batch
       real(c_double), value :: lr
       real(c_double), intent(out) :: loss_per_epoch(*)
     end function fgpu_ssh_nn_train

     integer(c_int) function fgpu_ssh_solve_learned(ctx, x, rhs, soltol, maxiter, fallback_maxiter, stats, &
          fallback_iters) bind(C, name="fgpu_ssh_solve_learned")
       import :: c_int, c_ptr, c_double, fgpu_cg_stats
       type(c_ptr), value :: ctx
       real(c_double), intent(inout) :: x(*)
       real(c_double), intent(in) :: rhs(*)
       real(c_double), value :: soltol
       integer(c_int), value :: maxiter, fallback_maxiter
       type(fgpu_cg_stats), intent(out) :: stats
       integer(c_int), intent(out) :: fallback_iters
     end function fgpu_ssh_solve_learned

     integer(c_int) function fgpu_ssh_solve_learned_resident(ctx, soltol, maxiter, fallback_maxiter, nrep, stats, &
          fallback_iters) bind(C, name="fgpu_ssh_solve_learned_resident")
       import :: c_int, c_ptr, c_double, fgpu_cg_stats
       type(c_ptr), value :: ctx
       real(c_double), value :: soltol
       integer(c_int), value :: maxiter, fallback_maxiter, nrep
       type(fgpu_cg_stats), intent(out) :: stats
       integer(c_int), intent(out) :: fallback_iters
     end function fgpu_ssh_solve_learned_resident
  end interface
end module fesom_gpu_c
