! fesom_gpu_host.f90 -- Fortran host side of the MI355X tracer-transport / SSH path.
!
! Mirrors the reference's call surface for one partition resident on one GPU:
!   init_tracers_AB(tr_num)                   oce_tracer_mod.F90:13-143
!   do_oce_adv_tra(dt, vel, w, wi, we, tr_num) oce_adv_tra_driver.F90:46-66
!   diff_tracers_ale(dt, tr_num)              oce_ale_tracer.F90:533-793 (+ :318-322)
!   diff_ver_part_impl_ale(dt, tr_num)        oce_ale_tracer.F90:907-1435
!   solve_tracers_ale(dt, nsteps)             oce_ale_tracer.F90:135-529 (tracer loop, batched)
!   init_stiff_mat_ale(dt, g, alpha, theta)   oce_ale.F90:1393-1582
!   ssh_solve_preconditioner(solverinfo)      solver.F90:31-117
!   ssh_solve_cg(x, rhs, solverinfo)          solver.F90:120-385
!   solve_ssh_ale(d_eta, ssh_rhs, solverinfo) oce_ale.F90:3056-3099
! The dynamics/tracers/mesh derived-type arguments of the reference are held by
! the device context of this module (uploaded once by gpu_mesh_setup,
! gpu_set_vertical, gpu_set_state and gpu_tracer_put); velocities passed to
! do_oce_adv_tra are uploaded on each call like the reference's arguments.
! A failing call prints the library's message and stops, as par_ex does for
! an unknown scheme or boundary-condition ID.
module fesom_gpu_host
  use, intrinsic :: iso_c_binding
  use fesom_gpu_c
  implicit none
  private

  integer, parameter, public :: WP = c_double

  ! mirrors T_SOLVERINFO (MOD_DYN.F90:13-46): knobs and per-run counters
  type, public :: t_gpu_solverinfo
     integer :: maxiter = 2000
     real(WP) :: soltol = 1e-5_WP
     integer :: precond_variant = 1
     integer :: iters_last = 0, iters_max = 0, iters_sum = 0, nsolves = 0
     integer :: nonconv = 0, nbreakdown = 0, nnegrz = 0
     real(WP) :: resid_last = 0.0_WP, rtol_last = 0.0_WP
     real(WP) :: ms_last = 0.0_WP
     logical :: pr_ready = .false.
     ! learned preconditioner: FCG cap before the plain-CG fallback, and the
     ! fallback's share of the last solve
     integer :: maxiter_learned = 500
     integer :: fallback_last = 0
  end type t_gpu_solverinfo

  type(c_ptr), save :: ctx = c_null_ptr

  ! valuesold as the reference holds it, (AB_order-1, nl-1, nod2D), for AB2
  ! and AB3 tracers; or, for an AB2 tracer, its one old level (nl-1, nod2D)
  interface gpu_tracer_put
     module procedure gpu_tracer_put3, gpu_tracer_put2
  end interface gpu_tracer_put
  interface gpu_tracer_get
     module procedure gpu_tracer_get3, gpu_tracer_get2
  end interface gpu_tracer_get

  public :: gpu_set_ki
  public :: gpu_init, gpu_finalize, gpu_check, gpu_mesh_setup, gpu_set_vertical, gpu_set_state, &
       gpu_tracers_init, gpu_tracer_config, gpu_tracer_put, gpu_tracer_get, gpu_tracer_content, &
       init_tracers_AB, do_oce_adv_tra, diff_tracers_ale, diff_ver_part_impl_ale, solve_tracers_ale, &
       init_stiff_mat_ale, ssh_solve_preconditioner, ssh_solve_cg, solve_ssh_ale, gpu_synchronize, &
       gpu_autotune_step, gpu_ssh_nn_init, gpu_ssh_nn_train, ssh_solve_cg_learned, vert_vel_ale, &
       gpu_vert_vel_ale, compute_ssh_rhs_ale, update_stiff_mat_ale, compute_hbar_ale, &
       gpu_set_ale_config

contains

  subroutine gpu_check(status, where)
    integer(c_int), intent(in) :: status
    character(len=*), intent(in) :: where
    character(kind=c_char), pointer :: msg(:)
    integer :: n
    if (status == FGPU_OK) return
    call c_f_pointer(fgpu_last_error(), msg, [4096])
    n = 0
    do while (n < 4096)
       if (msg(n + 1) == c_null_char) exit
       n = n + 1
    end do
    write(*, '(a,a,a,i0,a,4096a)') 'fesom_gpu: ', where, ' failed with status ', status, ': ', msg(1:n)
    error stop 1
  end subroutine gpu_check

  function cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: c
    c = trim(s) // c_null_char
  end function cstr

  subroutine gpu_init(device)
    integer, intent(in) :: device
    if (c_associated(ctx)) return
    call gpu_check(fgpu_create(ctx, int(device, c_int)), 'fgpu_create')
  end subroutine gpu_init

  subroutine gpu_finalize()
    if (.not. c_associated(ctx)) return
    call gpu_check(fgpu_destroy(ctx), 'fgpu_destroy')
    ctx = c_null_ptr
  end subroutine gpu_finalize

  subroutine gpu_synchronize()
    call gpu_check(fgpu_synchronize(ctx), 'fgpu_synchronize')
  end subroutine gpu_synchronize

  ! t_mesh components (MOD_MESH.F90:22-175), 1-based as the reference holds them
  subroutine gpu_mesh_setup(nod2D, elem2D, edge2D, nl, cyclic_length, coord_nod2D, elem2D_nodes, &
       edges, edge_tri, nod_in_elem2D, nod_in_elem2D_num, nlevels, ulevels, nlevels_nod2D, &
       ulevels_nod2D, nlevels_nod2D_min, ulevels_nod2D_max, elem_area, elem_cos, edge_dxdy, &
       edge_cross_dxdy, gradient_sca)
    integer, intent(in) :: nod2D, elem2D, edge2D, nl
    real(WP), intent(in) :: cyclic_length
    real(WP), intent(in) :: coord_nod2D(:, :), elem_area(:), elem_cos(:), edge_dxdy(:, :), &
         edge_cross_dxdy(:, :), gradient_sca(:, :)
    integer, intent(in) :: elem2D_nodes(:, :), edges(:, :), edge_tri(:, :), nod_in_elem2D(:, :), &
         nod_in_elem2D_num(:), nlevels(:), ulevels(:), nlevels_nod2D(:), ulevels_nod2D(:), &
         nlevels_nod2D_min(:), ulevels_nod2D_max(:)
    call gpu_check(fgpu_mesh_init(ctx, int(nod2D, c_int), int(elem2D, c_int), int(edge2D, c_int), &
         int(nl, c_int), int(size(nod_in_elem2D, 1), c_int), cyclic_length, coord_nod2D, elem2D_nodes, &
         edges, edge_tri, nod_in_elem2D, nod_in_elem2D_num, nlevels, ulevels, nlevels_nod2D, &
         ulevels_nod2D, nlevels_nod2D_min, ulevels_nod2D_max, elem_area, elem_cos, edge_dxdy, &
         edge_cross_dxdy, gradient_sca), 'fgpu_mesh_init')
  end subroutine gpu_mesh_setup

  subroutine gpu_set_vertical(area, areasvol, hnode, hnode_new, helem, zbar_3d_n, Z_3d_n, zbar_n_bot, &
       zbar_e_bot, zbar_e_srf)
    real(WP), intent(in) :: area(:, :), areasvol(:, :), hnode(:, :), hnode_new(:, :), helem(:, :), &
         zbar_3d_n(:, :), Z_3d_n(:, :), zbar_n_bot(:), zbar_e_bot(:), zbar_e_srf(:)
    call gpu_check(fgpu_set_vertical(ctx, area, areasvol, hnode, hnode_new, helem, zbar_3d_n, Z_3d_n, &
         zbar_n_bot, zbar_e_bot, zbar_e_srf), 'fgpu_set_vertical')
  end subroutine gpu_set_vertical

  ! velocities (t_dyn), o_ARRAYS%Kv and the bc_surface forcing arrays;
  ! use_wsplit is dynamics%use_wsplit (default .false.)
  subroutine gpu_set_state(uv, w, w_e, w_i, Kv, heat_flux, water_flux, virtual_salt, relax_salt, &
       real_salt_flux, is_nonlinfs, use_wsplit)
    real(WP), intent(in) :: uv(:, :, :), w(:, :), w_e(:, :), w_i(:, :), Kv(:, :)
    real(WP), intent(in) :: heat_flux(:), water_flux(:), virtual_salt(:), relax_salt(:), real_salt_flux(:)
    real(WP), intent(in) :: is_nonlinfs
    logical, intent(in), optional :: use_wsplit
    call gpu_check(fgpu_set_velocity(ctx, uv, w, w_e, w_i, wsplit_code(use_wsplit)), 'fgpu_set_velocity')
    call gpu_check(fgpu_set_kv(ctx, Kv), 'fgpu_set_kv')
    call gpu_check(fgpu_set_surface_forcing(ctx, heat_flux, water_flux, virtual_salt, relax_salt, &
         real_salt_flux, is_nonlinfs), 'fgpu_set_surface_forcing')
  end subroutine gpu_set_state

  ! o_ARRAYS Ki(nl-1,nod2D), slope_tapered(3,nl-1,nod2D) and Redi (oce_modules.F90:75):
  ! horizontal/Redi diffusion of diff_tracers_ale. Without Ki the diffusivity is zero.
  subroutine gpu_set_ki(Ki, slope_tapered, Redi)
    real(WP), intent(in), target, optional :: Ki(:, :), slope_tapered(:, :, :)
    logical, intent(in), optional :: Redi
    type(c_ptr) :: pk, ps
    integer(c_int) :: r
    pk = c_null_ptr
    ps = c_null_ptr
    r = 0_c_int
    if (present(Ki)) pk = c_loc(Ki)
    if (present(slope_tapered)) ps = c_loc(slope_tapered)
    if (present(Redi)) then
      if (Redi) r = 1_c_int
    end if
    call gpu_check(fgpu_set_ki(ctx, pk, ps, r), 'fgpu_set_ki')
  end subroutine gpu_set_ki

  integer(c_int) function wsplit_code(use_wsplit)
    logical, intent(in), optional :: use_wsplit
    wsplit_code = 0_c_int
    if (present(use_wsplit)) then
      if (use_wsplit) wsplit_code = 1_c_int
    end if
  end function wsplit_code

  integer(c_int) function ale_code(which_ALE)
    character(len=*), intent(in) :: which_ALE
    select case (trim(which_ALE))
    case ('linfs')
       ale_code = FGPU_ALE_LINFS
    case ('zlevel')
       ale_code = FGPU_ALE_ZLEVEL
    case ('zstar')
       ale_code = FGPU_ALE_ZSTAR
    case default
       ale_code = -1
    end select
  end function ale_code

  ! vert_vel_ale (oce_ale.F90:2107-2668) with compute_CFLz and compute_Wvel_split,
  ! on the device from uv; the results stay resident for the transport.
  ! hbar, hbar_old are read under zstar and zlevel (zlevel also needs gpu_set_ale_config).
  subroutine gpu_vert_vel_ale(dt, which_ALE, uv, hbar, hbar_old, use_wsplit, wsplit_maxcfl, cflmax)
    real(WP), intent(in) :: dt, wsplit_maxcfl
    character(len=*), intent(in) :: which_ALE
    real(WP), intent(in) :: uv(:, :, :), hbar(:), hbar_old(:)
    logical, intent(in) :: use_wsplit
    real(WP), intent(out) :: cflmax
    integer(c_int) :: nan_columns, ws
    ws = 0
    if (use_wsplit) ws = 1
    call gpu_check(fgpu_set_uv(ctx, uv), 'fgpu_set_uv')
    call gpu_check(fgpu_vert_vel_ale(ctx, dt, ale_code(which_ALE), hbar, hbar_old, ws, wsplit_maxcfl, cflmax, &
         nan_columns), 'vert_vel_ale')
    if (nan_columns > 0) write(*,*) ' --> subroutine vert_vel_ale --> found Nan in Wvel in', nan_columns, 'columns'
  end subroutine gpu_vert_vel_ale

  ! the same, with Wvel, Wvel_e, Wvel_i, CFL_z (nl,nod2D) and hnode_new
  ! (nl-1,nod2D) copied back as the reference's dynamics and mesh hold them
  subroutine vert_vel_ale(dt, which_ALE, uv, hbar, hbar_old, use_wsplit, wsplit_maxcfl, Wvel, Wvel_e, Wvel_i, &
       CFL_z, hnode_new)
    real(WP), intent(in) :: dt, wsplit_maxcfl
    character(len=*), intent(in) :: which_ALE
    real(WP), intent(in) :: uv(:, :, :), hbar(:), hbar_old(:)
    logical, intent(in) :: use_wsplit
    real(WP), intent(out) :: Wvel(:, :), Wvel_e(:, :), Wvel_i(:, :), CFL_z(:, :), hnode_new(:, :)
    real(WP) :: cflmax
    call gpu_vert_vel_ale(dt, which_ALE, uv, hbar, hbar_old, use_wsplit, wsplit_maxcfl, cflmax)
    call gpu_check(fgpu_vert_vel_download(ctx, Wvel, Wvel_e, Wvel_i, CFL_z, hnode_new), 'fgpu_vert_vel_download')
  end subroutine vert_vel_ale

  ! compute_ssh_rhs_ale (oce_ale.F90:1803-1938) on the device from uv and
  ! uv_rhs; ssh_rhs also becomes the SSH solve's right-hand side
  subroutine compute_ssh_rhs_ale(uv, uv_rhs, ssh_rhs_old, alpha, which_ALE, use_cavity_fw2press, ssh_rhs)
    real(WP), intent(in) :: uv(:, :, :), uv_rhs(:, :, :), ssh_rhs_old(:), alpha
    character(len=*), intent(in) :: which_ALE
    logical, intent(in) :: use_cavity_fw2press
    real(WP), intent(out) :: ssh_rhs(:)
    integer(c_int) :: cav
    cav = 0
    if (use_cavity_fw2press) cav = 1
    call gpu_check(fgpu_set_uv(ctx, uv), 'fgpu_set_uv')
    call gpu_check(fgpu_compute_ssh_rhs_ale(ctx, uv_rhs, ssh_rhs_old, alpha, ale_code(which_ALE), cav, ssh_rhs), &
         'compute_ssh_rhs_ale')
  end subroutine compute_ssh_rhs_ale

  ! which_ALE = 'zlevel' settings: mesh%zbar, min_hnode, lzstar_lev (namelist ale_def)
  subroutine gpu_set_ale_config(zbar, min_hnode, lzstar_lev)
    real(WP), intent(in) :: zbar(:), min_hnode
    integer, intent(in) :: lzstar_lev
    call gpu_check(fgpu_set_ale_config(ctx, zbar, min_hnode, int(lzstar_lev, c_int)), 'fgpu_set_ale_config')
  end subroutine gpu_set_ale_config

  ! compute_hbar_ale (oce_ale.F90:1956-2091) from uv: hbar is updated in place,
  ! hbar_old, ssh_rhs_old and dhe are written, like mesh%hbar, hbar_old and
  ! dynamics%ssh_rhs_old, mesh%dhe in the reference
  subroutine compute_hbar_ale(dt, which_ALE, uv, hbar, hbar_old, ssh_rhs_old, dhe)
    real(WP), intent(in) :: dt, uv(:, :, :)
    character(len=*), intent(in) :: which_ALE
    real(WP), intent(inout) :: hbar(:)
    real(WP), intent(out) :: hbar_old(:), ssh_rhs_old(:), dhe(:)
    real(WP), allocatable :: h0(:)
    h0 = hbar
    call gpu_check(fgpu_set_uv(ctx, uv), 'fgpu_set_uv')
    call gpu_check(fgpu_compute_hbar_ale(ctx, dt, ale_code(which_ALE), h0, hbar, hbar_old, ssh_rhs_old, dhe), &
         'compute_hbar_ale')
  end subroutine compute_hbar_ale

  ! update_stiff_mat_ale (oce_ale.F90:1683-1792): the stiffness term of the
  ! element ssh change dhe(elem2D) added to the assembled SSH matrix
  subroutine update_stiff_mat_ale(dhe, dt, g, alpha, theta)
    real(WP), intent(in) :: dhe(:), dt, g, alpha, theta
    call gpu_check(fgpu_update_stiff_mat_ale(ctx, dhe, dt, g, alpha, theta), 'update_stiff_mat_ale')
  end subroutine update_stiff_mat_ale

  subroutine gpu_tracers_init(num_tracers)
    integer, intent(in) :: num_tracers
    call gpu_check(fgpu_tracers_init(ctx, int(num_tracers, c_int)), 'fgpu_tracers_init')
  end subroutine gpu_tracers_init

  ! one tracer's namelist entry (config/namelist.tra:33-38)
  subroutine gpu_tracer_config(tr_num, ID, AB_order, tra_adv_hor, tra_adv_ver, tra_adv_lim, &
       tra_adv_ph, tra_adv_pv, i_vert_diff)
    integer, intent(in) :: tr_num, ID, AB_order
    character(len=*), intent(in) :: tra_adv_hor, tra_adv_ver, tra_adv_lim
    real(WP), intent(in) :: tra_adv_ph, tra_adv_pv
    logical, intent(in) :: i_vert_diff
    integer(c_int) :: ivd
    ivd = merge(1_c_int, 0_c_int, i_vert_diff)
    call gpu_check(fgpu_tracer_config(ctx, int(tr_num, c_int), int(ID, c_int), int(AB_order, c_int), &
         cstr(tra_adv_hor), cstr(tra_adv_ver), cstr(tra_adv_lim), tra_adv_ph, tra_adv_pv, ivd), &
         'fgpu_tracer_config')
  end subroutine gpu_tracer_config

  ! tracers%data(tr_num)%values(nl-1,nod2D) and %valuesold(AB_order-1,nl-1,nod2D)
  subroutine gpu_tracer_put3(tr_num, values, valuesold)
    integer, intent(in) :: tr_num
    real(WP), intent(in) :: values(:, :), valuesold(:, :, :)
    call gpu_check(fgpu_tracer_upload(ctx, int(tr_num, c_int), values, valuesold), 'fgpu_tracer_upload')
  end subroutine gpu_tracer_put3

  subroutine gpu_tracer_put2(tr_num, values, valuesold)
    integer, intent(in) :: tr_num
    real(WP), intent(in) :: values(:, :), valuesold(:, :)
    call gpu_check(fgpu_tracer_upload(ctx, int(tr_num, c_int), values, valuesold), 'fgpu_tracer_upload')
  end subroutine gpu_tracer_put2

  subroutine gpu_tracer_get3(tr_num, values, valuesold)
    integer, intent(in) :: tr_num
    real(WP), intent(out) :: values(:, :), valuesold(:, :, :)
    call gpu_check(fgpu_tracer_download(ctx, int(tr_num, c_int), values, valuesold), 'fgpu_tracer_download')
  end subroutine gpu_tracer_get3

  subroutine gpu_tracer_get2(tr_num, values, valuesold)
    integer, intent(in) :: tr_num
    real(WP), intent(out) :: values(:, :), valuesold(:, :)
    call gpu_check(fgpu_tracer_download(ctx, int(tr_num, c_int), values, valuesold), 'fgpu_tracer_download')
  end subroutine gpu_tracer_get2

  real(WP) function gpu_tracer_content(tr_num)
    integer, intent(in) :: tr_num
    call gpu_check(fgpu_tracer_content(ctx, int(tr_num, c_int), gpu_tracer_content), 'fgpu_tracer_content')
  end function gpu_tracer_content

  ! ---- reference call surface ------------------------------------------------

  subroutine init_tracers_AB(tr_num)
    integer, intent(in) :: tr_num
    call gpu_check(fgpu_init_tracers_AB(ctx, int(tr_num, c_int)), 'init_tracers_AB')
  end subroutine init_tracers_AB

  ! vel(2,nl-1,elem2D); W, WI, WE(nl,nod2D). Leaves del_ttf_advhoriz/advvert on
  ! the device for diff_tracers_ale, like tracers%work in the reference.
  ! use_wsplit is dynamics%use_wsplit, which the reference reads inside
  ! (oce_adv_tra_driver.F90:323); default .false.
  subroutine do_oce_adv_tra(dt, vel, w, wi, we, tr_num, use_wsplit)
    real(WP), intent(in) :: dt
    real(WP), intent(in) :: vel(:, :, :), w(:, :), wi(:, :), we(:, :)
    integer, intent(in) :: tr_num
    logical, intent(in), optional :: use_wsplit
    call gpu_check(fgpu_set_velocity(ctx, vel, w, we, wi, wsplit_code(use_wsplit)), 'do_oce_adv_tra (velocity)')
    call gpu_check(fgpu_do_oce_adv_tra(ctx, dt, int(tr_num, c_int)), 'do_oce_adv_tra')
  end subroutine do_oce_adv_tra

  subroutine diff_tracers_ale(dt, tr_num)
    real(WP), intent(in) :: dt
    integer, intent(in) :: tr_num
    call gpu_check(fgpu_diff_tracers_ale(ctx, dt, int(tr_num, c_int)), 'diff_tracers_ale')
  end subroutine diff_tracers_ale

  subroutine diff_ver_part_impl_ale(dt, tr_num)
    real(WP), intent(in) :: dt
    integer, intent(in) :: tr_num
    call gpu_check(fgpu_diff_ver_part_impl_ale(ctx, dt, int(tr_num, c_int)), 'diff_ver_part_impl_ale')
  end subroutine diff_ver_part_impl_ale

  ! every tracer advances together (AB, advection, T*, vertical diffusion,
  ! salinity clamp); nsteps > 1 replays the captured step
  subroutine solve_tracers_ale(dt, nsteps)
    real(WP), intent(in) :: dt
    integer, intent(in) :: nsteps
    call gpu_check(fgpu_solve_tracers_ale(ctx, dt, int(nsteps, c_int)), 'solve_tracers_ale')
  end subroutine solve_tracers_ale

  subroutine init_stiff_mat_ale(dt, g, alpha, theta)
    real(WP), intent(in) :: dt, g, alpha, theta
    call gpu_check(fgpu_init_stiff_mat_ale(ctx, dt, g, alpha, theta), 'init_stiff_mat_ale')
  end subroutine init_stiff_mat_ale

  subroutine ssh_solve_preconditioner(solverinfo)
    type(t_gpu_solverinfo), intent(inout) :: solverinfo
    call gpu_check(fgpu_ssh_solve_preconditioner(ctx, int(solverinfo%precond_variant, c_int)), &
         'ssh_solve_preconditioner')
    solverinfo%pr_ready = .true.
  end subroutine ssh_solve_preconditioner

  ! x(nod2D) in/out (warm start); counters and rate-limited warnings as in
  ! solver.F90:319-380. CG never stops the run.
  subroutine ssh_solve_cg(x, rhs, solverinfo)
    real(WP), intent(inout) :: x(:)
    real(WP), intent(in) :: rhs(:)
    type(t_gpu_solverinfo), intent(inout) :: solverinfo
    type(fgpu_cg_stats) :: st
    call gpu_check(fgpu_ssh_solve_cg(ctx, x, rhs, solverinfo%soltol, int(solverinfo%maxiter, c_int), st), &
         'ssh_solve_cg')
    if (st%nnegrz > 0 .and. solverinfo%nnegrz == 0) &
         write(*, *) 'WARNING: ssh CG r.z < 0 -- preconditioner is not symmetric positive definite'
    solverinfo%nnegrz = solverinfo%nnegrz + st%nnegrz
    if (st%nbreakdown > 0) then
       solverinfo%nbreakdown = solverinfo%nbreakdown + st%nbreakdown
       if (solverinfo%nbreakdown <= 5) write(*, *) 'WARNING: ssh CG breakdown: previous r.z is zero'
    end if
    solverinfo%iters_last = st%iters_last
    solverinfo%iters_max = max(solverinfo%iters_max, st%iters_last)
    solverinfo%iters_sum = solverinfo%iters_sum + st%iters_last
    solverinfo%nsolves = solverinfo%nsolves + 1
    solverinfo%resid_last = st%resid_last
    solverinfo%rtol_last = st%rtol_last
    solverinfo%ms_last = st%ms
    if (st%converged == 0) then
       solverinfo%nonconv = solverinfo%nonconv + 1
       if (solverinfo%nonconv <= 5 .or. mod(solverinfo%nonconv, 100) == 0) &
            write(*, *) 'WARNING: ssh CG did not converge in ', solverinfo%maxiter, ' iters; rms(resid)=', &
            solverinfo%resid_last, ' target rtol=', solverinfo%rtol_last, ' (occurrence ', solverinfo%nonconv, ')'
    end if
  end subroutine ssh_solve_cg

  ! lazy preconditioner build, then the CG solve of d_eta (oce_ale.F90:3094-3097);
  ! the halo exchange is the one-partition no-op
  subroutine solve_ssh_ale(d_eta, ssh_rhs, solverinfo)
    real(WP), intent(inout) :: d_eta(:)
    real(WP), intent(in) :: ssh_rhs(:)
    type(t_gpu_solverinfo), intent(inout) :: solverinfo
    if (.not. solverinfo%pr_ready) call ssh_solve_preconditioner(solverinfo)
    call ssh_solve_cg(d_eta, ssh_rhs, solverinfo)
  end subroutine solve_ssh_ale

  ! fastest launch shape of the batched step for this partition (tracer state kept)
  subroutine gpu_autotune_step(dt, nsteps, tracers_per_block, ms_per_step)
    real(WP), intent(in) :: dt
    integer, intent(in) :: nsteps
    integer, intent(out) :: tracers_per_block
    real(WP), intent(out) :: ms_per_step
    integer(c_int) :: tpb
    call gpu_check(fgpu_autotune_step(ctx, dt, int(nsteps, c_int), tpb, ms_per_step), 'fgpu_autotune_step')
    tracers_per_block = tpb
  end subroutine gpu_autotune_step

  subroutine gpu_ssh_nn_init(seed)
    integer, intent(in) :: seed
    call gpu_check(fgpu_ssh_nn_init(ctx, int(seed, c_long_long)), 'fgpu_ssh_nn_init')
  end subroutine gpu_ssh_nn_init

  ! b(nod2D, nsamp) right-hand sides with solutions x(nod2D, nsamp)
  subroutine gpu_ssh_nn_train(b, x, epochs, batch, lr, loss_per_epoch)
    real(WP), intent(in) :: b(:, :), x(:, :)
    integer, intent(in) :: epochs, batch
    real(WP), intent(in) :: lr
    real(WP), intent(out) :: loss_per_epoch(:)
    call gpu_check(fgpu_ssh_nn_train(ctx, b, x, int(size(b, 2), c_int), int(epochs, c_int), int(batch, c_int), lr, &
         loss_per_epoch), 'fgpu_ssh_nn_train')
  end subroutine gpu_ssh_nn_train

  ! ssh_solve_cg with the learned preconditioner in flexible CG; plain CG
  ! continues from its x when it breaks down or reaches maxiter_learned
  subroutine ssh_solve_cg_learned(x, rhs, solverinfo)
    real(WP), intent(inout) :: x(:)
    real(WP), intent(in) :: rhs(:)
    type(t_gpu_solverinfo), intent(inout) :: solverinfo
    type(fgpu_cg_stats) :: st
    integer(c_int) :: fb
    call gpu_check(fgpu_ssh_solve_learned(ctx, x, rhs, solverinfo%soltol, int(solverinfo%maxiter_learned, c_int), &
         int(solverinfo%maxiter, c_int), st, fb), 'ssh_solve_cg_learned')
    solverinfo%iters_last = st%iters_last
    solverinfo%iters_max = max(solverinfo%iters_max, st%iters_last)
    solverinfo%iters_sum = solverinfo%iters_sum + st%iters_last
    solverinfo%nsolves = solverinfo%nsolves + 1
    solverinfo%resid_last = st%resid_last
    solverinfo%rtol_last = st%rtol_last
    solverinfo%ms_last = st%ms
    solverinfo%fallback_last = fb
    solverinfo%nbreakdown = solverinfo%nbreakdown + st%nbreakdown
    if (st%converged == 0) solverinfo%nonconv = solverinfo%nonconv + 1
  end subroutine ssh_solve_cg_learned

end module fesom_gpu_host
