! fgpu_dropin_driver.f90 -- runs a case through the drop-in entry points with
! the reference's argument lists (fesom_gpu_dropin.f90), on the reference's
! derived types (fesom_types.f90):
!   * nsingle steps of the per-tracer body of solve_tracers_ale
!     (oce_ale_tracer.F90:280-385): init_tracers_AB on the host
!     (oce_tracer_mod.F90:44-59, 94-101), then
!       call do_oce_adv_tra(dt, dynamics%uv, dynamics%w, dynamics%w_i, dynamics%w_e, tr_num, &
!                           dynamics, tracers, partit, mesh)          (oce_ale_tracer.F90:312)
!     the del_ttf sum (:318-322) and the T* update of diff_tracers_ale
!     (:742-755) on the host, then
!       call diff_ver_part_impl_ale(tr_num, dynamics, tracers, ice, partit, mesh)   (:772)
!     and the salinity clamp (:502-514);
!   * nbatched steps of call solve_tracers_ale(ice, dynamics, tracers, partit, mesh);
!   * call solve_ssh_ale(dynamics, partit, mesh) on the case's ssh_rhs with the
!     stiffness matrix of the matrix file (d_eta starts at zero).
!
!   fgpu_dropin_driver <case.bin> <matrix.bin> <result.bin>
!
! case.bin is the case file of fgpu_case_driver (format 1, AB2 tracers, no Ki).
! matrix.bin: int32 dim, nza, rowptr(dim+1), colind(nza) (1-based, diagonal
! first in each row); real values(nza).
! result.bin: int32 status(8): iters_last, converged, nbreakdown, nnegrz,
! nonconv, precond_variant, 0, 0; real resid_last, rtol_last;
! values(nl-1,N2,ntr); valuesold(nl-1,N2,ntr); d_eta(N2); pr_values(nza).
program fgpu_dropin_driver
  use, intrinsic :: iso_c_binding
  use o_PARAM
  use MOD_PARTIT
  use MOD_MESH
  use MOD_TRACER
  use MOD_DYN
  use MOD_ICE
  use o_ARRAYS
  use g_forcing_arrays, only: real_salt_flux
  use g_config, only: dt, cyclic_length
  use oce_adv_tra_driver_interfaces
  use diff_ver_part_impl_ale_interface
  use solve_tracers_ale_interface
  use oce_ale_interfaces
  use fesom_gpu_dropin_ctx, only: dropin_finalize
  implicit none

  integer, parameter :: MAGIC = 1179076949
  real(WP), parameter :: epsilon = 0.1_WP   ! oce_modules.F90:105
  integer(c_int32_t) :: hdr(16), ostat(8), mdim, mnza
  real(WP) :: rhdr(4)
  integer :: N2, F2, E2, nl, kmax, ntr, nsingle, nbatched, istep, tr, u, n, nz, i
  character(len=4096) :: fin, fmat, fout
  type(t_partit), target :: partit
  type(t_mesh), target :: mesh
  type(t_tracer), target :: tracers
  type(t_dyn), target :: dynamics
  type(t_ice), target :: ice
  integer(c_int32_t), allocatable :: ids(:)
  real(WP), allocatable :: values(:, :, :), valuesold(:, :, :)

  call get_command_argument(1, fin)
  call get_command_argument(2, fmat)
  call get_command_argument(3, fout)
  if (len_trim(fin) == 0 .or. len_trim(fmat) == 0 .or. len_trim(fout) == 0) then
     write(*, *) 'usage: fgpu_dropin_driver <case.bin> <matrix.bin> <result.bin>'
     error stop 2
  end if

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(u) hdr
  if (hdr(1) /= MAGIC .or. hdr(2) /= 1 .or. hdr(13) /= 0 .or. hdr(16) /= 2) then
     write(*, *) 'fgpu_dropin_driver: not a case file of format 1 with AB2 tracers and no Ki: ', trim(fin)
     error stop 2
  end if
  N2 = hdr(3); F2 = hdr(4); E2 = hdr(5); nl = hdr(6); kmax = hdr(7); ntr = hdr(8)
  nsingle = hdr(9); nbatched = hdr(10)
  read(u) rhdr
  dt = rhdr(1); cyclic_length = rhdr(2)
  dynamics%solverinfo%soltol = rhdr(3)
  dynamics%solverinfo%precond_variant = hdr(11)
  dynamics%solverinfo%maxiter = hdr(12)
  dynamics%use_wsplit = hdr(15) == 1

  partit%myDim_nod2D = N2
  partit%myDim_elem2D = F2
  partit%myDim_edge2D = E2
  mesh%nod2D = N2; mesh%elem2D = F2; mesh%edge2D = E2; mesh%nl = nl
  allocate(mesh%coord_nod2D(2, N2), mesh%elem2D_nodes(3, F2), mesh%edges(2, E2), mesh%edge_tri(2, E2), &
       mesh%nod_in_elem2D(kmax, N2), mesh%nod_in_elem2D_num(N2), mesh%nlevels(F2), mesh%ulevels(F2), &
       mesh%nlevels_nod2D(N2), mesh%ulevels_nod2D(N2), mesh%nlevels_nod2D_min(N2), mesh%ulevels_nod2D_max(N2), &
       mesh%elem_area(F2), mesh%elem_cos(F2), mesh%edge_dxdy(2, E2), mesh%edge_cross_dxdy(4, E2), &
       mesh%gradient_sca(6, F2))
  read(u) mesh%coord_nod2D, mesh%elem2D_nodes, mesh%edges, mesh%edge_tri, mesh%nod_in_elem2D, &
       mesh%nod_in_elem2D_num, mesh%nlevels, mesh%ulevels, mesh%nlevels_nod2D, mesh%ulevels_nod2D, &
       mesh%nlevels_nod2D_min, mesh%ulevels_nod2D_max, mesh%elem_area, mesh%elem_cos, mesh%edge_dxdy, &
       mesh%edge_cross_dxdy, mesh%gradient_sca
  allocate(mesh%area(nl, N2), mesh%areasvol(nl, N2), mesh%hnode(nl - 1, N2), mesh%hnode_new(nl - 1, N2), &
       mesh%helem(nl - 1, F2), mesh%zbar_3d_n(nl, N2), mesh%Z_3d_n(nl - 1, N2), mesh%zbar_n_bot(N2), &
       mesh%zbar_e_bot(F2), mesh%zbar_e_srf(F2))
  read(u) mesh%area, mesh%areasvol, mesh%hnode, mesh%hnode_new, mesh%helem, mesh%zbar_3d_n, mesh%Z_3d_n, &
       mesh%zbar_n_bot, mesh%zbar_e_bot, mesh%zbar_e_srf
  allocate(dynamics%uv(2, nl - 1, F2), dynamics%w(nl, N2), dynamics%w_e(nl, N2), dynamics%w_i(nl, N2), Kv(nl, N2))
  read(u) dynamics%uv, dynamics%w, dynamics%w_e, dynamics%w_i, Kv
  allocate(heat_flux(N2), water_flux(N2), virtual_salt(N2), relax_salt(N2), real_salt_flux(N2))
  read(u) heat_flux, water_flux, virtual_salt, relax_salt, real_salt_flux
  allocate(ids(ntr), values(nl - 1, N2, ntr), valuesold(nl - 1, N2, ntr), dynamics%ssh_rhs(N2), dynamics%d_eta(N2))
  read(u) ids, values, valuesold, dynamics%ssh_rhs
  close(u)
  is_nonlinfs = 0.0_WP

  ! tracers as tracer_init leaves them (oce_setup_step.F90:470-480), with the
  ! namelist entry MFCT / QR4C / FCT, ph = 0, pv = 1 (config/namelist.tra:35-36)
  tracers%num_tracers = ntr
  allocate(tracers%data(ntr))
  do tr = 1, ntr
     associate (d => tracers%data(tr))
       allocate(d%values(nl - 1, N2), d%valuesAB(nl - 1, N2), d%valuesold(1, nl - 1, N2))
       d%values = values(:, :, tr)
       d%valuesold(1, :, :) = valuesold(:, :, tr)
       d%valuesAB = 0.0_WP
       d%ID = ids(tr)
       d%AB_order = 2
       d%tra_adv_hor = 'MFCT'
       d%tra_adv_ver = 'QR4C'
       d%tra_adv_lim = 'FCT'
       d%tra_adv_ph = 0.0_WP
       d%tra_adv_pv = 1.0_WP
       d%i_vert_diff = .true.
       d%ltra_diag = .false.
     end associate
  end do
  allocate(tracers%work%del_ttf(nl - 1, N2), tracers%work%del_ttf_advhoriz(nl - 1, N2), &
       tracers%work%del_ttf_advvert(nl - 1, N2))
  allocate(ice%data(1))
  allocate(ice%data(1)%values(N2))
  ice%data(1)%values = 0.0_WP

  ! the stiffness matrix init_stiff_mat_ale assembled (one partition: local = global numbering)
  open(newunit=u, file=trim(fmat), access='stream', form='unformatted', status='old', action='read')
  read(u) mdim, mnza
  mesh%ssh_stiff%dim = mdim
  mesh%ssh_stiff%nza = mnza
  allocate(mesh%ssh_stiff%rowptr(mdim + 1), mesh%ssh_stiff%colind(mnza), mesh%ssh_stiff%values(mnza))
  read(u) mesh%ssh_stiff%rowptr, mesh%ssh_stiff%colind, mesh%ssh_stiff%values
  close(u)
  mesh%ssh_stiff%rowptr_loc = mesh%ssh_stiff%rowptr
  mesh%ssh_stiff%colind_loc = mesh%ssh_stiff%colind

  ! ---- the per-tracer step through do_oce_adv_tra and diff_ver_part_impl_ale
  do istep = 1, nsingle
     do tr = 1, ntr
        associate (d => tracers%data(tr), wk => tracers%work)
          wk%del_ttf = 0.0_WP
          wk%del_ttf_advhoriz = 0.0_WP
          wk%del_ttf_advvert = 0.0_WP
          d%valuesAB = -(0.5_WP + epsilon) * d%valuesold(1, :, :) + (1.5_WP + epsilon) * d%values
          d%valuesold(1, :, :) = d%values
          call do_oce_adv_tra(dt, dynamics%uv, dynamics%w, dynamics%w_i, dynamics%w_e, tr, dynamics, tracers, &
               partit, mesh)
          wk%del_ttf = wk%del_ttf + wk%del_ttf_advhoriz + wk%del_ttf_advvert
          do n = 1, N2
             do nz = mesh%ulevels_nod2D(n), mesh%nlevels_nod2D(n) - 1
                wk%del_ttf(nz, n) = wk%del_ttf(nz, n) + d%values(nz, n) * (mesh%hnode(nz, n) - mesh%hnode_new(nz, n))
                d%values(nz, n) = d%values(nz, n) + wk%del_ttf(nz, n) / mesh%hnode_new(nz, n)
             end do
          end do
        end associate
        call diff_ver_part_impl_ale(tr, dynamics, tracers, ice, partit, mesh)
     end do
     if (ntr >= 2) then
        do n = 1, N2
           do nz = mesh%ulevels_nod2D(n), mesh%nlevels_nod2D(n) - 1
              if (tracers%data(2)%values(nz, n) > 45.0_WP) tracers%data(2)%values(nz, n) = 45.0_WP
              if (tracers%data(2)%values(nz, n) < 3.0_WP) tracers%data(2)%values(nz, n) = 3.0_WP
           end do
        end do
     end if
  end do

  ! ---- the batched step
  do istep = 1, nbatched
     call solve_tracers_ale(ice, dynamics, tracers, partit, mesh)
  end do

  ! ---- the SSH solve
  dynamics%d_eta = 0.0_WP
  call solve_ssh_ale(dynamics, partit, mesh)

  ostat = 0
  ostat(1) = dynamics%solverinfo%iters_last
  ostat(2) = merge(1, 0, dynamics%solverinfo%nonconv == 0)
  ostat(3) = dynamics%solverinfo%nbreakdown
  ostat(4) = dynamics%solverinfo%nnegrz
  ostat(5) = dynamics%solverinfo%nonconv
  ostat(6) = dynamics%solverinfo%precond_variant
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write(u) ostat
  write(u) dynamics%solverinfo%resid_last, dynamics%solverinfo%rtol_last
  write(u) (tracers%data(i)%values, i = 1, ntr)
  write(u) (tracers%data(i)%valuesold(1, :, :), i = 1, ntr)
  write(u) dynamics%d_eta
  write(u) mesh%ssh_stiff%pr_values
  close(u)
  call dropin_finalize()
end program fgpu_dropin_driver
