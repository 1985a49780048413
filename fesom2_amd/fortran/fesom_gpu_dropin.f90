! fesom_gpu_dropin.f90 -- the reference's entry points of the tracer transport,
! the implicit vertical diffusion and the SSH solve, with the reference's own
! argument lists, as external subroutines that run on the GPU library:
!
!   do_oce_adv_tra(dt, vel, w, wi, we, tr_num, dynamics, tracers, partit, mesh)
!                                                    oce_adv_tra_driver.F90:3-19, 46-66
!   diff_ver_part_impl_ale(tr_num, dynamics, tracers, ice, partit, mesh)
!                                                    oce_ale_tracer.F90:44-62, 907-1435
!   solve_tracers_ale(ice, dynamics, tracers, partit, mesh)
!                                                    oce_ale_tracer.F90:114-130, 135-529
!   ssh_solve_preconditioner(solverinfo, partit, mesh)   solver.F90:1-13, 31-117
!   ssh_solve_cg(x, rhs, solverinfo, partit, mesh)       solver.F90:15-29, 120-385
!   solve_ssh_ale(dynamics, partit, mesh)                oce_ale.F90:106-114, 3056-3099
!
! and the interface modules under which the reference's call sites see them
! (oce_adv_tra_driver_interfaces, diff_ver_part_impl_ale_interface,
! solve_tracers_ale_interface, ssh_solve_preconditioner_interface,
! ssh_solve_cg_interface, oce_ale_interfaces), so that call sites such as
! oce_ale_tracer.F90:312,772 and oce_ale.F90:3660 compile unchanged.
!
! Data flow. The mesh connectivity is uploaded on the first call (and again if
! the mesh changes size); the 3-D geometry, the velocities, the module arrays
! the reference reads (o_ARRAYS Kv, Ki, slope_tapered and the surface fluxes,
! g_forcing_arrays) and the tracer they work on are uploaded on every call,
! because the reference reads them from the derived types on every call, and
! the results are written back into the same components the reference writes:
!   do_oce_adv_tra         tracers%work%del_ttf_advhoriz, %del_ttf_advvert
!   diff_ver_part_impl_ale tracers%data(tr_num)%values
!   solve_tracers_ale      every tracer's values and valuesold (batched step)
!   ssh_solve_preconditioner  mesh%ssh_stiff%pr_values
!   ssh_solve_cg           x and the solverinfo counters
! do_oce_adv_tra takes valuesAB from tracers%data(tr_num) as the caller's
! init_tracers_AB left it and recomputes the element gradients of values, as
! that routine does. ssh_solve_cg solves with mesh%ssh_stiff%values and
! %pr_values as they are at the call.
!
! Errors: a failing library call prints its message and stops the run, as
! par_ex does in the reference (unknown scheme, unknown bc_surface ID). The
! CG records nonconv, nbreakdown and nnegrz and never stops the run.
!
! The device is the one the environment variable FESOM_GPU_DEVICE names
! (default 0).
module fesom_gpu_dropin_ctx
  use, intrinsic :: iso_c_binding
  use fesom_gpu_c
  use o_PARAM, only: WP, Redi
  use MOD_MESH
  use MOD_TRACER
  use MOD_DYN
  implicit none
  private

  type(c_ptr), save :: dctx = c_null_ptr
  integer, save :: mesh_key(6) = -1
  integer, save :: ntr_dev = -1
  ! per tracer: the configuration last handed to the device
  integer, save, allocatable :: cfg_int(:, :)
  character(len=60), save, allocatable :: cfg_str(:)
  real(WP), save, allocatable :: cfg_real(:, :)

  public :: dctx, dropin_check, dropin_mesh, dropin_vertical, dropin_tracers, dropin_tracer_put, &
       dropin_module_arrays, dropin_ssh_matrix, dropin_finalize

contains

  subroutine dropin_check(status, where)
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
    write(*, *) 'the model will stop!'
    error stop 1
  end subroutine dropin_check

  function cstr(s) result(c)
    character(len=*), intent(in) :: s
    character(kind=c_char, len=len_trim(s) + 1) :: c
    c = trim(s) // c_null_char
  end function cstr

  subroutine dropin_finalize()
    if (.not. c_associated(dctx)) return
    call dropin_check(fgpu_destroy(dctx), 'fgpu_destroy')
    dctx = c_null_ptr
    mesh_key = -1
    ntr_dev = -1
  end subroutine dropin_finalize

  ! context on first use; the mesh connectivity whenever the mesh differs in size
  subroutine dropin_mesh(mesh)
    use g_config, only: cyclic_length
    type(t_mesh), intent(in) :: mesh
    integer :: key(6), dev, ios
    character(len=16) :: env
    if (.not. c_associated(dctx)) then
       dev = 0
       call get_environment_variable('FESOM_GPU_DEVICE', env)
       if (len_trim(env) > 0) then
          read(env, *, iostat=ios) dev
          if (ios /= 0) dev = 0
       end if
       call dropin_check(fgpu_create(dctx, int(dev, c_int)), 'fgpu_create')
    end if
    key = [mesh%nod2D, mesh%elem2D, mesh%edge2D, mesh%nl, size(mesh%nod_in_elem2D, 1), &
         size(mesh%ssh_stiff%values)]
    if (all(key(1:5) == mesh_key(1:5))) return
    call dropin_check(fgpu_mesh_init(dctx, int(mesh%nod2D, c_int), int(mesh%elem2D, c_int), &
         int(mesh%edge2D, c_int), int(mesh%nl, c_int), int(size(mesh%nod_in_elem2D, 1), c_int), &
         cyclic_length, mesh%coord_nod2D, mesh%elem2D_nodes, mesh%edges, mesh%edge_tri, mesh%nod_in_elem2D, &
         mesh%nod_in_elem2D_num, mesh%nlevels, mesh%ulevels, mesh%nlevels_nod2D, mesh%ulevels_nod2D, &
         mesh%nlevels_nod2D_min, mesh%ulevels_nod2D_max, mesh%elem_area, mesh%elem_cos, mesh%edge_dxdy, &
         mesh%edge_cross_dxdy, mesh%gradient_sca), 'mesh_setup (fgpu_mesh_init)')
    mesh_key = key
    ntr_dev = -1
  end subroutine dropin_mesh

  ! the 3-D geometry of the current ALE state, as the reference reads it from mesh
  subroutine dropin_vertical(mesh)
    type(t_mesh), intent(in) :: mesh
    call dropin_check(fgpu_set_vertical(dctx, mesh%area, mesh%areasvol, mesh%hnode, mesh%hnode_new, mesh%helem, &
         mesh%zbar_3d_n, mesh%Z_3d_n, mesh%zbar_n_bot, mesh%zbar_e_bot, mesh%zbar_e_srf), 'fgpu_set_vertical')
  end subroutine dropin_vertical

  ! tracer slots and namelist entries (tra_adv_hor/ver/lim, ph, pv, AB_order, ID,
  ! i_vert_diff) of every tracer; only what changed is handed over again
  subroutine dropin_tracers(tracers)
    type(t_tracer), intent(in) :: tracers
    integer :: i, ci(3)
    character(len=60) :: cs
    real(WP) :: cr(2)
    if (tracers%num_tracers /= ntr_dev) then
       call dropin_check(fgpu_tracers_init(dctx, int(tracers%num_tracers, c_int)), 'fgpu_tracers_init')
       ntr_dev = tracers%num_tracers
       if (allocated(cfg_int)) deallocate(cfg_int, cfg_str, cfg_real)
       allocate(cfg_int(3, ntr_dev), cfg_str(ntr_dev), cfg_real(2, ntr_dev))
       cfg_int = -1
    end if
    do i = 1, tracers%num_tracers
       associate (d => tracers%data(i))
         ci = [d%ID, d%AB_order, merge(1, 0, d%i_vert_diff)]
         cs = d%tra_adv_hor // d%tra_adv_ver // d%tra_adv_lim
         cr = [d%tra_adv_ph, d%tra_adv_pv]
         if (all(ci == cfg_int(:, i)) .and. cs == cfg_str(i) .and. all(cr == cfg_real(:, i))) cycle
         call dropin_check(fgpu_tracer_config(dctx, int(i, c_int), int(d%ID, c_int), int(d%AB_order, c_int), &
              cstr(d%tra_adv_hor), cstr(d%tra_adv_ver), cstr(d%tra_adv_lim), d%tra_adv_ph, d%tra_adv_pv, &
              int(ci(3), c_int)), 'tracer configuration (fgpu_tracer_config)')
         cfg_int(:, i) = ci
         cfg_str(i) = cs
         cfg_real(:, i) = cr
       end associate
    end do
  end subroutine dropin_tracers

  subroutine dropin_tracer_put(tracers, tr_num)
    type(t_tracer), intent(in) :: tracers
    integer, intent(in) :: tr_num
    call dropin_check(fgpu_tracer_upload(dctx, int(tr_num, c_int), tracers%data(tr_num)%values, &
         tracers%data(tr_num)%valuesold), 'fgpu_tracer_upload')
  end subroutine dropin_tracer_put

  ! module arrays the vertical diffusion reads: o_ARRAYS Kv, the bc_surface
  ! fluxes (heat_flux, water_flux, virtual_salt, relax_salt; real_salt_flux and
  ! prec_rain of g_forcing_arrays; wiso_flux_oce, hosing_flux/3D), Ki and
  ! slope_tapered with Redi
  subroutine dropin_module_arrays(nod2D)
    use o_ARRAYS
    use g_forcing_arrays
    use g_config, only: use_hosing, hosing_mode
    integer, intent(in) :: nod2D
    real(WP), allocatable, target :: zero(:)
    ! target copies of the module arrays whose addresses the library takes
    real(WP), allocatable, target :: t_rain(:), t_wiso(:, :), t_hos(:), t_hos3(:, :), t_ki(:, :), t_sl(:, :, :)
    type(c_ptr) :: p_rain, p_wiso, p_hos, p_hos3
    integer(c_int) :: depth
    if (.not. allocated(Kv)) then
       write(*, *) 'fesom_gpu: o_ARRAYS Kv is not allocated'
       error stop 1
    end if
    call dropin_check(fgpu_set_kv(dctx, Kv), 'fgpu_set_kv')
    allocate(zero(nod2D))
    zero = 0.0_WP
    call dropin_check(fgpu_set_surface_forcing(dctx, or_zero(heat_flux), or_zero(water_flux), &
         or_zero(virtual_salt), or_zero(relax_salt), or_zero(real_salt_flux), is_nonlinfs), &
         'fgpu_set_surface_forcing')
    p_rain = c_null_ptr
    p_wiso = c_null_ptr
    p_hos = c_null_ptr
    p_hos3 = c_null_ptr
    if (allocated(prec_rain)) then
       t_rain = prec_rain
       p_rain = c_loc(t_rain)
    end if
    if (allocated(wiso_flux_oce)) then
       t_wiso = wiso_flux_oce
       p_wiso = c_loc(t_wiso)
    end if
    if (allocated(hosing_flux)) then
       t_hos = hosing_flux
       p_hos = c_loc(t_hos)
    end if
    if (allocated(hosing_flux3D)) then
       t_hos3 = hosing_flux3D
       p_hos3 = c_loc(t_hos3)
    end if
    depth = 0
    if (use_hosing .and. trim(hosing_mode) == 'depth') depth = 1
    if (c_associated(p_rain) .or. c_associated(p_wiso) .or. c_associated(p_hos) .or. c_associated(p_hos3)) &
         call dropin_check(fgpu_set_surface_forcing_tracers(dctx, p_rain, p_wiso, p_hos, p_hos3, depth), &
         'fgpu_set_surface_forcing_tracers')
    if (allocated(Ki)) then
       t_ki = Ki
       if (Redi .and. allocated(slope_tapered)) then
          t_sl = slope_tapered
          call dropin_check(fgpu_set_ki(dctx, c_loc(t_ki), c_loc(t_sl), 1_c_int), 'fgpu_set_ki')
       else
          call dropin_check(fgpu_set_ki(dctx, c_loc(t_ki), c_null_ptr, 0_c_int), 'fgpu_set_ki')
       end if
    else
       call dropin_check(fgpu_set_ki(dctx, c_null_ptr, c_null_ptr, 0_c_int), 'fgpu_set_ki')
    end if
  contains
    ! a flux array the run never allocated reads as zero, as an unused one does
    function or_zero(a) result(p)
      real(WP), intent(in), allocatable, target :: a(:)
      real(WP), pointer :: p(:)
      if (allocated(a)) then
         p => a
      else
         p => zero
      end if
    end function or_zero
  end subroutine dropin_module_arrays

  ! mesh%ssh_stiff as assembled by init_stiff_mat_ale / update_stiff_mat_ale
  ! (local numbering rowptr_loc / colind_loc, diagonal first in each row)
  subroutine dropin_ssh_matrix(mesh)
    type(t_mesh), intent(in) :: mesh
    integer :: n
    n = mesh%ssh_stiff%dim
    if (n <= 0) n = size(mesh%ssh_stiff%rowptr_loc) - 1
    call dropin_check(fgpu_ssh_set_matrix(dctx, int(n, c_int), int(size(mesh%ssh_stiff%values), c_int), &
         mesh%ssh_stiff%rowptr_loc, mesh%ssh_stiff%colind_loc, mesh%ssh_stiff%values), &
         'SSH matrix (fgpu_ssh_set_matrix)')
  end subroutine dropin_ssh_matrix

end module fesom_gpu_dropin_ctx

! ------------------------------------------------------------------------------
! interface modules of the reference's call sites
! ------------------------------------------------------------------------------
module oce_adv_tra_driver_interfaces
  interface
     subroutine do_oce_adv_tra(dt, vel, w, wi, we, tr_num, dynamics, tracers, partit, mesh)
       use MOD_MESH
       use MOD_TRACER
       use MOD_PARTIT
       use MOD_PARSUP
       use MOD_DYN
       real(kind=WP),  intent(in),    target :: dt
       integer,        intent(in)            :: tr_num
       type(t_partit), intent(inout), target :: partit
       type(t_mesh)  , intent(in)   , target :: mesh
       type(t_tracer), intent(inout), target :: tracers
       type(t_dyn)   , intent(inout), target :: dynamics
       real(kind=WP),  intent(in)            :: vel(2, mesh%nl-1, partit%myDim_elem2D+partit%eDim_elem2D)
       real(kind=WP),  intent(in), target    :: W(mesh%nl,    partit%myDim_nod2D+partit%eDim_nod2D)
       real(kind=WP),  intent(in), target    :: WI(mesh%nl,   partit%myDim_nod2D+partit%eDim_nod2D)
       real(kind=WP),  intent(in), target    :: WE(mesh%nl,   partit%myDim_nod2D+partit%eDim_nod2D)
     end subroutine do_oce_adv_tra
  end interface
end module oce_adv_tra_driver_interfaces

module diff_ver_part_impl_ale_interface
  interface
     subroutine diff_ver_part_impl_ale(tr_num, dynamics, tracer, ice, partit, mesh)
       use MOD_MESH
       use MOD_PARTIT
       use MOD_PARSUP
       use MOD_TRACER
       use MOD_DYN
       use MOD_ICE
       integer       , intent(in)   , target :: tr_num
       type(t_dyn)   , intent(inout), target :: dynamics
       type(t_tracer), intent(inout), target :: tracer
       type(t_partit), intent(inout), target :: partit
       type(t_mesh)  , intent(in)   , target :: mesh
       type(t_ice)   , intent(in)   , target :: ice
     end subroutine diff_ver_part_impl_ale
  end interface
end module diff_ver_part_impl_ale_interface

module solve_tracers_ale_interface
  interface
     subroutine solve_tracers_ale(ice, dynamics, tracers, partit, mesh)
       use MOD_MESH
       use MOD_PARTIT
       use MOD_PARSUP
       use MOD_TRACER
       use MOD_DYN
       use MOD_ICE
       type(t_ice)   , intent(in),    target :: ice
       type(t_dyn)   , intent(inout), target :: dynamics
       type(t_tracer), intent(inout), target :: tracers
       type(t_partit), intent(inout), target :: partit
       type(t_mesh)  , intent(in)   , target :: mesh
     end subroutine solve_tracers_ale
  end interface
end module solve_tracers_ale_interface

module ssh_solve_preconditioner_interface
  interface
     subroutine ssh_solve_preconditioner(solverinfo, partit, mesh)
       use MOD_MESH
       use MOD_PARTIT
       use MOD_PARSUP
       use MOD_DYN
       type(t_solverinfo),  intent(inout), target :: solverinfo
       type(t_partit),      intent(inout), target :: partit
       type(t_mesh),        intent(inout), target :: mesh
     end subroutine ssh_solve_preconditioner
  end interface
end module ssh_solve_preconditioner_interface

module ssh_solve_cg_interface
  interface
     subroutine ssh_solve_cg(x, rhs, solverinfo, partit, mesh)
       use MOD_MESH
       use MOD_PARTIT
       use MOD_PARSUP
       use MOD_DYN
       type(t_solverinfo),  intent(inout), target :: solverinfo
       type(t_partit),      intent(inout), target :: partit
       type(t_mesh),        intent(inout), target :: mesh
       real(kind=WP),       intent(inout) :: x(partit%myDim_nod2D+partit%eDim_nod2D)
       real(kind=WP),       intent(in)    :: rhs(partit%myDim_nod2D+partit%eDim_nod2D)
     end subroutine ssh_solve_cg
  end interface
end module ssh_solve_cg_interface

module oce_ale_interfaces
  interface
     subroutine solve_ssh_ale(dynamics, partit, mesh)
       use MOD_MESH
       use MOD_PARTIT
       use MOD_PARSUP
       use MOD_DYN
       type(t_dyn)   , intent(inout), target :: dynamics
       type(t_partit), intent(inout), target :: partit
       type(t_mesh)  , intent(inout), target :: mesh
     end subroutine solve_ssh_ale
  end interface
end module oce_ale_interfaces

! ------------------------------------------------------------------------------
! the entry points
! ------------------------------------------------------------------------------
subroutine do_oce_adv_tra(dt, vel, w, wi, we, tr_num, dynamics, tracers, partit, mesh)
  use, intrinsic :: iso_c_binding
  use fesom_gpu_c
  use MOD_MESH
  use MOD_TRACER
  use MOD_PARTIT
  use MOD_PARSUP
  use MOD_DYN
  use fesom_gpu_dropin_ctx
  implicit none
  real(kind=WP),  intent(in),    target :: dt
  integer,        intent(in)            :: tr_num
  type(t_mesh)  , intent(in)   , target :: mesh
  type(t_partit), intent(inout), target :: partit
  type(t_tracer), intent(inout), target :: tracers
  type(t_dyn)   , intent(inout), target :: dynamics
  real(kind=WP),  intent(in)            :: vel(2, mesh%nl-1, partit%myDim_elem2D+partit%eDim_elem2D)
  real(kind=WP),  intent(in), target    :: W(mesh%nl,    partit%myDim_nod2D+partit%eDim_nod2D)
  real(kind=WP),  intent(in), target    :: WI(mesh%nl,   partit%myDim_nod2D+partit%eDim_nod2D)
  real(kind=WP),  intent(in), target    :: WE(mesh%nl,   partit%myDim_nod2D+partit%eDim_nod2D)
  integer :: n2, l
  call dropin_mesh(mesh)
  call dropin_vertical(mesh)
  call dropin_check(fgpu_set_velocity(dctx, vel, W, WE, WI, merge(1_c_int, 0_c_int, dynamics%use_wsplit)), &
       'do_oce_adv_tra (velocities)')
  call dropin_tracers(tracers)
  call dropin_tracer_put(tracers, tr_num)
  call dropin_check(fgpu_tracer_set_valuesAB(dctx, int(tr_num, c_int), tracers%data(tr_num)%valuesAB), &
       'do_oce_adv_tra (valuesAB)')
  call dropin_check(fgpu_do_oce_adv_tra(dctx, dt, int(tr_num, c_int)), 'do_oce_adv_tra')
  n2 = partit%myDim_nod2D + partit%eDim_nod2D
  l = mesh%nl - 1
  if (.not. allocated(tracers%work%del_ttf_advhoriz)) allocate(tracers%work%del_ttf_advhoriz(l, n2))
  if (.not. allocated(tracers%work%del_ttf_advvert)) allocate(tracers%work%del_ttf_advvert(l, n2))
  call dropin_check(fgpu_tracer_work_download(dctx, int(tr_num, c_int), 'del_ttf_advhoriz' // c_null_char, &
       tracers%work%del_ttf_advhoriz, int(size(tracers%work%del_ttf_advhoriz), c_long)), 'do_oce_adv_tra (del_ttf)')
  call dropin_check(fgpu_tracer_work_download(dctx, int(tr_num, c_int), 'del_ttf_advvert' // c_null_char, &
       tracers%work%del_ttf_advvert, int(size(tracers%work%del_ttf_advvert), c_long)), 'do_oce_adv_tra (del_ttf)')
end subroutine do_oce_adv_tra

subroutine diff_ver_part_impl_ale(tr_num, dynamics, tracer, ice, partit, mesh)
  use, intrinsic :: iso_c_binding
  use fesom_gpu_c
  use MOD_MESH
  use MOD_PARTIT
  use MOD_PARSUP
  use MOD_TRACER
  use MOD_DYN
  use MOD_ICE
  use g_config, only: dt
  use fesom_gpu_dropin_ctx
  implicit none
  integer       , intent(in)   , target :: tr_num
  type(t_dyn)   , intent(inout), target :: dynamics
  type(t_tracer), intent(inout), target :: tracer
  type(t_partit), intent(inout), target :: partit
  type(t_mesh)  , intent(in)   , target :: mesh
  type(t_ice)   , intent(in)   , target :: ice
  call dropin_mesh(mesh)
  call dropin_vertical(mesh)
  ! w_i of the split vertical velocity enters the rows of non-FCT tracers
  call dropin_check(fgpu_set_velocity(dctx, dynamics%uv, dynamics%w, dynamics%w_e, dynamics%w_i, &
       merge(1_c_int, 0_c_int, dynamics%use_wsplit)), 'diff_ver_part_impl_ale (velocities)')
  call dropin_module_arrays(mesh%nod2D)
  call dropin_tracers(tracer)
  call dropin_tracer_put(tracer, tr_num)
  call dropin_check(fgpu_diff_ver_part_impl_ale(dctx, dt, int(tr_num, c_int)), 'diff_ver_part_impl_ale')
  call dropin_check(fgpu_tracer_download(dctx, int(tr_num, c_int), tracer%data(tr_num)%values, &
       tracer%data(tr_num)%valuesold), 'diff_ver_part_impl_ale (values)')
end subroutine diff_ver_part_impl_ale

! The whole tracer step of solve_tracers_ale for every tracer at once (AB,
! advection, T*, vertical diffusion, salinity clamp) in the fused batched
! launches, from and back to the tracers' values and valuesold.
subroutine solve_tracers_ale(ice, dynamics, tracers, partit, mesh)
  use, intrinsic :: iso_c_binding
  use fesom_gpu_c
  use MOD_MESH
  use MOD_PARTIT
  use MOD_PARSUP
  use MOD_TRACER
  use MOD_DYN
  use MOD_ICE
  use g_config, only: dt
  use fesom_gpu_dropin_ctx
  implicit none
  type(t_ice)   , intent(in),    target :: ice
  type(t_dyn)   , intent(inout), target :: dynamics
  type(t_tracer), intent(inout), target :: tracers
  type(t_partit), intent(inout), target :: partit
  type(t_mesh)  , intent(in)   , target :: mesh
  integer :: i
  call dropin_mesh(mesh)
  call dropin_vertical(mesh)
  call dropin_check(fgpu_set_velocity(dctx, dynamics%uv, dynamics%w, dynamics%w_e, dynamics%w_i, &
       merge(1_c_int, 0_c_int, dynamics%use_wsplit)), 'solve_tracers_ale (velocities)')
  call dropin_module_arrays(mesh%nod2D)
  call dropin_tracers(tracers)
  do i = 1, tracers%num_tracers
     call dropin_tracer_put(tracers, i)
  end do
  call dropin_check(fgpu_solve_tracers_ale(dctx, dt, 1_c_int), 'solve_tracers_ale')
  do i = 1, tracers%num_tracers
     call dropin_check(fgpu_tracer_download(dctx, int(i, c_int), tracers%data(i)%values, &
          tracers%data(i)%valuesold), 'solve_tracers_ale (values)')
  end do
end subroutine solve_tracers_ale

subroutine ssh_solve_preconditioner(solverinfo, partit, mesh)
  use, intrinsic :: iso_c_binding
  use fesom_gpu_c
  use MOD_MESH
  use MOD_PARTIT
  use MOD_PARSUP
  use MOD_DYN
  use fesom_gpu_dropin_ctx
  implicit none
  type(t_solverinfo),  intent(inout), target :: solverinfo
  type(t_partit),      intent(inout), target :: partit
  type(t_mesh),        intent(inout), target :: mesh
  integer(c_int) :: nza, nza_out
  call dropin_mesh(mesh)
  call dropin_vertical(mesh)
  call dropin_ssh_matrix(mesh)
  call dropin_check(fgpu_ssh_solve_preconditioner(dctx, int(solverinfo%precond_variant, c_int)), &
       'ssh_solve_preconditioner')
  nza = int(size(mesh%ssh_stiff%values), c_int)
  if (.not. allocated(mesh%ssh_stiff%pr_values)) allocate(mesh%ssh_stiff%pr_values(nza))
  call dropin_check(fgpu_ssh_get_matrix(dctx, c_null_ptr, c_null_ptr, c_null_ptr, c_loc(mesh%ssh_stiff%pr_values), &
       nza_out), 'ssh_solve_preconditioner (pr_values)')
end subroutine ssh_solve_preconditioner

! x(nod2D) in/out (warm start). Counters and rate-limited warnings as in
! solver.F90:319-380; a solve that does not converge never stops the run.
subroutine ssh_solve_cg(x, rhs, solverinfo, partit, mesh)
  use, intrinsic :: iso_c_binding
  use fesom_gpu_c
  use MOD_MESH
  use MOD_PARTIT
  use MOD_PARSUP
  use MOD_DYN
  use fesom_gpu_dropin_ctx
  implicit none
  type(t_solverinfo),  intent(inout), target :: solverinfo
  type(t_partit),      intent(inout), target :: partit
  type(t_mesh),        intent(inout), target :: mesh
  real(kind=WP),       intent(inout) :: x(partit%myDim_nod2D+partit%eDim_nod2D)
  real(kind=WP),       intent(in)    :: rhs(partit%myDim_nod2D+partit%eDim_nod2D)
  type(fgpu_cg_stats) :: st
  call dropin_mesh(mesh)
  call dropin_vertical(mesh)
  call dropin_ssh_matrix(mesh)
  if (.not. allocated(mesh%ssh_stiff%pr_values)) then
     write(*, *) 'fesom_gpu: ssh_solve_cg needs mesh%ssh_stiff%pr_values (ssh_solve_preconditioner)'
     error stop 1
  end if
  call dropin_check(fgpu_ssh_set_preconditioner(dctx, int(solverinfo%precond_variant, c_int), &
       mesh%ssh_stiff%pr_values), 'ssh_solve_cg (preconditioner)')
  call dropin_check(fgpu_ssh_solve_cg(dctx, x, rhs, solverinfo%soltol, int(solverinfo%maxiter, c_int), st), &
       'ssh_solve_cg')
  if (st%nnegrz > 0) then
     if (solverinfo%nnegrz == 0) &
          write(*, *) 'WARNING: ssh CG r.z < 0 -- preconditioner is not symmetric positive definite'
     solverinfo%nnegrz = solverinfo%nnegrz + st%nnegrz
  end if
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
  if (st%converged == 0) then
     solverinfo%nonconv = solverinfo%nonconv + 1
     if (solverinfo%nonconv <= 5 .or. mod(solverinfo%nonconv, 100) == 0) &
          write(*, *) 'WARNING: ssh CG did not converge in ', solverinfo%maxiter, ' iters; rms(resid)=', &
          solverinfo%resid_last, ' target rtol=', solverinfo%rtol_last, ' (occurrence ', solverinfo%nonconv, ')'
  end if
end subroutine ssh_solve_cg

! preconditioner built when mesh%ssh_stiff%pr_values is not yet allocated,
! then the CG solve of d_eta (oce_ale.F90:3088-3097); the halo exchange that
! follows is the one-partition no-op
subroutine solve_ssh_ale(dynamics, partit, mesh)
  use MOD_MESH
  use MOD_PARTIT
  use MOD_PARSUP
  use MOD_DYN
  use ssh_solve_preconditioner_interface
  use ssh_solve_cg_interface
  implicit none
  type(t_dyn)   , intent(inout), target :: dynamics
  type(t_partit), intent(inout), target :: partit
  type(t_mesh)  , intent(inout), target :: mesh
  if (.not. allocated(mesh%ssh_stiff%pr_values)) &
       call ssh_solve_preconditioner(dynamics%solverinfo, partit, mesh)
  call ssh_solve_cg(dynamics%d_eta, dynamics%ssh_rhs, dynamics%solverinfo, partit, mesh)
end subroutine solve_ssh_ale
