! fgpu_case_driver.f90 -- Fortran host program over fesom_gpu_host.
!
! Reads one partition's mesh, state and tracers from a stream file, advances
! the tracers through the reference-shaped per-tracer sequence
! (init_tracers_AB, do_oce_adv_tra, diff_tracers_ale; oce_ale_tracer.F90:280-385)
! for nsingle steps and through the batched solve_tracers_ale for nbatched more,
! runs vert_vel_ale (linfs) and compute_ssh_rhs_ale (alpha = 1/2, uv_rhs = uv,
! ssh_rhs_old = the case's ssh_rhs) on the transported state's velocity,
! assembles the SSH operator and solves it with solve_ssh_ale, then writes the
! results to a second stream file.
!
!   fgpu_case_driver <case.bin> <result.bin>
!   fgpu_case_driver --abi        (prints the library version; no GPU use)
!   fgpu_case_driver --read <case.bin>   (reads the case, prints sizes and sums; no GPU use)
!
! Case file (little-endian, Fortran array order, integers int32, reals FP64):
!   int32 header(16): magic 1179076949, format 1, nod2D, elem2D, edge2D, nl, kmax,
!                     ntr, nsingle, nbatched, precond_variant, maxiter, has_Ki, Redi,
!                     use_wsplit, 0
!   real  dt, cyclic_length, soltol, g
!   coord_nod2D(2,N2) elem2D_nodes(3,F2) edges(2,E2) edge_tri(2,E2)
!   nod_in_elem2D(kmax,N2) nod_in_elem2D_num(N2) nlevels(F2) ulevels(F2)
!   nlevels_nod2D(N2) ulevels_nod2D(N2) nlevels_nod2D_min(N2) ulevels_nod2D_max(N2)
!   elem_area(F2) elem_cos(F2) edge_dxdy(2,E2) edge_cross_dxdy(4,E2) gradient_sca(6,F2)
!   area(nl,N2) areasvol(nl,N2) hnode(nl-1,N2) hnode_new(nl-1,N2) helem(nl-1,F2)
!   zbar_3d_n(nl,N2) Z_3d_n(nl-1,N2) zbar_n_bot(N2) zbar_e_bot(F2) zbar_e_srf(F2)
!   uv(2,nl-1,F2) w(nl,N2) w_e(nl,N2) w_i(nl,N2) Kv(nl,N2)
!   heat_flux(N2) water_flux(N2) virtual_salt(N2) relax_salt(N2) real_salt_flux(N2)
!   ID(ntr) values(nl-1,N2,ntr) valuesold(nl-1,N2,ntr) ssh_rhs(N2)
!   [Ki(nl-1,N2) if has_Ki] [slope_tapered(3,nl-1,N2) if Redi]
!   [older(nl-1,N2) if header word 16 is 3: tracer 1 runs AB3 with old levels
!    valuesold(:,:,1) and older]
! Result file:
!   int32 status(8): iters_last, converged (1/0), nbreakdown, nnegrz, nonconv, 0, 0, 0
!   real  resid_last, rtol_last, content_before(ntr), content_after(ntr)
!   values(nl-1,N2,ntr) valuesold(nl-1,N2,ntr) d_eta(N2)
!   Wvel(nl,N2) Wvel_e(nl,N2) Wvel_i(nl,N2) CFL_z(nl,N2) ssh_rhs_dyn(N2)
!   [older(nl-1,N2) if tracer 1 runs AB3]
program fgpu_case_driver
  use, intrinsic :: iso_c_binding
  use fesom_gpu_c, only: fgpu_version
  use fesom_gpu_host
  implicit none

  integer, parameter :: MAGIC = 1179076949
  integer(c_int32_t) :: hdr(16), ostat(8)
  real(WP) :: rhdr(4)
  integer :: N2, F2, E2, nl, kmax, ntr, nsingle, nbatched, istep, tr, u
  logical :: read_only
  real(WP) :: dt, cyclic_length, g
  character(len=4096) :: fin, fout
  real(WP), allocatable :: coord_nod2D(:, :), elem_area(:), elem_cos(:), edge_dxdy(:, :), &
       edge_cross_dxdy(:, :), gradient_sca(:, :)
  integer(c_int32_t), allocatable :: elem2D_nodes(:, :), edges(:, :), edge_tri(:, :), nod_in_elem2D(:, :), &
       nod_in_elem2D_num(:), nlevels(:), ulevels(:), nlevels_nod2D(:), ulevels_nod2D(:), &
       nlevels_nod2D_min(:), ulevels_nod2D_max(:), ids(:)
  real(WP), allocatable :: area(:, :), areasvol(:, :), hnode(:, :), hnode_new(:, :), helem(:, :), &
       zbar_3d_n(:, :), Z_3d_n(:, :), zbar_n_bot(:), zbar_e_bot(:), zbar_e_srf(:)
  real(WP), allocatable :: uv(:, :, :), w(:, :), w_e(:, :), w_i(:, :), Kv(:, :)
  real(WP), allocatable :: heat_flux(:), water_flux(:), virtual_salt(:), relax_salt(:), real_salt_flux(:)
  real(WP), allocatable :: values(:, :, :), valuesold(:, :, :), ssh_rhs(:), d_eta(:), c0(:), c1(:)
  real(WP), allocatable :: wv(:, :), wv_e(:, :), wv_i(:, :), cfl_z(:, :), hn_new(:, :), rhs_dyn(:), hbar(:)
  real(WP) :: cflmax
  real(WP), allocatable, target :: Ki(:, :), slope_tapered(:, :, :)
  real(WP), allocatable :: older(:, :), vo3(:, :, :)
  integer :: ab1
  logical :: use_wsplit, redi
  type(t_gpu_solverinfo) :: si

  call get_command_argument(1, fin)
  if (trim(fin) == '--abi') then
     write(*, '(a,i0)') 'fesom_gpu library version ', fgpu_version()
     stop
  end if
  read_only = trim(fin) == '--read'
  if (read_only) then
     call get_command_argument(2, fin)
  else
     call get_command_argument(2, fout)
  end if
  if (len_trim(fin) == 0 .or. (.not. read_only .and. len_trim(fout) == 0)) then
     write(*, *) 'usage: fgpu_case_driver <case.bin> <result.bin> | --read <case.bin> | --abi'
     error stop 2
  end if

  open(newunit=u, file=trim(fin), access='stream', form='unformatted', status='old', action='read')
  read(u) hdr
  if (hdr(1) /= MAGIC .or. hdr(2) /= 1) then
     write(*, *) 'fgpu_case_driver: not a case file (format 1): ', trim(fin)
     error stop 2
  end if
  N2 = hdr(3); F2 = hdr(4); E2 = hdr(5); nl = hdr(6); kmax = hdr(7); ntr = hdr(8)
  nsingle = hdr(9); nbatched = hdr(10)
  read(u) rhdr
  dt = rhdr(1); cyclic_length = rhdr(2); g = rhdr(4)
  si%soltol = rhdr(3); si%precond_variant = hdr(11); si%maxiter = hdr(12)

  allocate(coord_nod2D(2, N2), elem2D_nodes(3, F2), edges(2, E2), edge_tri(2, E2), nod_in_elem2D(kmax, N2), &
       nod_in_elem2D_num(N2), nlevels(F2), ulevels(F2), nlevels_nod2D(N2), ulevels_nod2D(N2), &
       nlevels_nod2D_min(N2), ulevels_nod2D_max(N2), elem_area(F2), elem_cos(F2), edge_dxdy(2, E2), &
       edge_cross_dxdy(4, E2), gradient_sca(6, F2))
  read(u) coord_nod2D, elem2D_nodes, edges, edge_tri, nod_in_elem2D, nod_in_elem2D_num, nlevels, ulevels, &
       nlevels_nod2D, ulevels_nod2D, nlevels_nod2D_min, ulevels_nod2D_max, elem_area, elem_cos, edge_dxdy, &
       edge_cross_dxdy, gradient_sca
  allocate(area(nl, N2), areasvol(nl, N2), hnode(nl - 1, N2), hnode_new(nl - 1, N2), helem(nl - 1, F2), &
       zbar_3d_n(nl, N2), Z_3d_n(nl - 1, N2), zbar_n_bot(N2), zbar_e_bot(F2), zbar_e_srf(F2))
  read(u) area, areasvol, hnode, hnode_new, helem, zbar_3d_n, Z_3d_n, zbar_n_bot, zbar_e_bot, zbar_e_srf
  allocate(uv(2, nl - 1, F2), w(nl, N2), w_e(nl, N2), w_i(nl, N2), Kv(nl, N2))
  read(u) uv, w, w_e, w_i, Kv
  allocate(heat_flux(N2), water_flux(N2), virtual_salt(N2), relax_salt(N2), real_salt_flux(N2))
  read(u) heat_flux, water_flux, virtual_salt, relax_salt, real_salt_flux
  allocate(ids(ntr), values(nl - 1, N2, ntr), valuesold(nl - 1, N2, ntr), ssh_rhs(N2), d_eta(N2), c0(ntr), c1(ntr))
  read(u) ids, values, valuesold, ssh_rhs
  redi = hdr(14) == 1
  use_wsplit = hdr(15) == 1
  if (hdr(13) == 1) then
     allocate(Ki(nl - 1, N2))
     read(u) Ki
  end if
  if (redi) then
     allocate(slope_tapered(3, nl - 1, N2))
     read(u) slope_tapered
  end if
  ab1 = merge(3, 2, hdr(16) == 3)
  if (ab1 == 3) then
     allocate(older(nl - 1, N2))
     read(u) older
  end if
  close(u)
  write(*, '(a,i0,a,i0,a,i0,a,i0,a,i0)') '[driver] nod2D=', N2, ' elem2D=', F2, ' edge2D=', E2, ' nl=', nl, &
       ' ntr=', ntr
  flush(6)
  if (read_only) then
     write(*, '(a,i0,1x,i0,1x,i0)') 'sums_int ', sum(int(edges, 8)), sum(int(nod_in_elem2D, 8)), sum(int(ids, 8))
     write(*, '(a,3es25.16)') 'sums_real ', sum(values), sum(uv), sum(ssh_rhs)
     write(*, '(a,2es25.16)') 'last ', gradient_sca(6, F2), Kv(nl, N2)
     write(*, '(a,i0,1x,i0,1x,i0,1x,i0)') 'flags ', hdr(13), hdr(14), hdr(15), ab1
     if (allocated(Ki)) write(*, '(a,es25.16)') 'ki ', sum(Ki)
     if (allocated(slope_tapered)) write(*, '(a,es25.16)') 'slope ', slope_tapered(3, nl - 1, N2)
     stop
  end if

  ! partition set-up (mesh_setup / tracer_init / dynamics_init on the device)
  call gpu_init(0)
  call gpu_mesh_setup(N2, F2, E2, nl, cyclic_length, coord_nod2D, elem2D_nodes, edges, edge_tri, &
       nod_in_elem2D, nod_in_elem2D_num, nlevels, ulevels, nlevels_nod2D, ulevels_nod2D, nlevels_nod2D_min, &
       ulevels_nod2D_max, elem_area, elem_cos, edge_dxdy, edge_cross_dxdy, gradient_sca)
  call gpu_set_vertical(area, areasvol, hnode, hnode_new, helem, zbar_3d_n, Z_3d_n, zbar_n_bot, zbar_e_bot, &
       zbar_e_srf)
  call gpu_set_state(uv, w, w_e, w_i, Kv, heat_flux, water_flux, virtual_salt, relax_salt, real_salt_flux, &
       0.0_WP, use_wsplit)
  ! o_ARRAYS Ki and slope_tapered, Redi (horizontal and isoneutral diffusion)
  if (allocated(Ki)) then
     if (redi) then
        call gpu_set_ki(Ki, slope_tapered, .true.)
     else
        call gpu_set_ki(Ki)
     end if
  end if
  call gpu_tracers_init(ntr)
  do tr = 1, ntr
     if (tr == 1 .and. ab1 == 3) then
        ! the reference's valuesold(AB_order-1, nl-1, nod2D), newer level first
        allocate(vo3(2, nl - 1, N2))
        vo3(1, :, :) = valuesold(:, :, 1)
        vo3(2, :, :) = older
        call gpu_tracer_config(tr, ids(tr), 3, 'MFCT', 'QR4C', 'FCT', 0.0_WP, 1.0_WP, .true.)
        call gpu_tracer_put(tr, values(:, :, tr), vo3)
     else
        call gpu_tracer_config(tr, ids(tr), 2, 'MFCT', 'QR4C', 'FCT', 0.0_WP, 1.0_WP, .true.)
        call gpu_tracer_put(tr, values(:, :, tr), valuesold(:, :, tr))
     end if
     c0(tr) = gpu_tracer_content(tr)
  end do

  ! reference-shaped tracer loop (oce_ale_tracer.F90:280-385)
  do istep = 1, nsingle
     do tr = 1, ntr
        call init_tracers_AB(tr)
        call do_oce_adv_tra(dt, uv, w, w_i, w_e, tr, use_wsplit)
        call diff_tracers_ale(dt, tr)
     end do
  end do
  write(*, '(a,i0,a)') '[driver] ', nsingle, ' per-tracer steps done'
  flush(6)
  ! batched step
  if (nbatched > 0) call solve_tracers_ale(dt, nbatched)
  call gpu_synchronize()
  write(*, '(a,i0,a)') '[driver] ', nbatched, ' batched steps done'
  flush(6)
  do tr = 1, ntr
     if (tr == 1 .and. ab1 == 3) then
        call gpu_tracer_get(tr, values(:, :, tr), vo3)
        valuesold(:, :, 1) = vo3(1, :, :)
        older = vo3(2, :, :)
     else
        call gpu_tracer_get(tr, values(:, :, tr), valuesold(:, :, tr))
     end if
     c1(tr) = gpu_tracer_content(tr)
  end do

  ! the steps either side of the transport (oce_ale.F90:1803-1938, 2107-3049)
  allocate(wv(nl, N2), wv_e(nl, N2), wv_i(nl, N2), cfl_z(nl, N2), hn_new(nl - 1, N2), rhs_dyn(N2), hbar(N2))
  hbar = 0.0_WP
  call vert_vel_ale(dt, 'linfs', uv, hbar, hbar, .false., 1.0_WP, wv, wv_e, wv_i, cfl_z, hn_new)
  call compute_ssh_rhs_ale(uv, uv, ssh_rhs, 0.5_WP, 'linfs', .false., rhs_dyn)
  write(*, '(a,es10.3)') '[driver] vert_vel_ale max |W|=', maxval(abs(wv))
  flush(6)

  ! SSH: operator, lazy preconditioner, CG from d_eta = 0
  call init_stiff_mat_ale(dt, g, 1.0_WP, 1.0_WP)
  d_eta = 0.0_WP
  call solve_ssh_ale(d_eta, ssh_rhs, si)
  write(*, '(a,i0,a,es10.3,a,es10.3)') '[driver] ssh iters=', si%iters_last, ' resid=', si%resid_last, &
       ' rtol=', si%rtol_last
  flush(6)
  call gpu_finalize()

  ostat = 0
  ostat(1) = si%iters_last
  ostat(2) = merge(1, 0, si%nonconv == 0)
  ostat(3) = si%nbreakdown
  ostat(4) = si%nnegrz
  ostat(5) = si%nonconv
  open(newunit=u, file=trim(fout), access='stream', form='unformatted', status='replace', action='write')
  write(u) ostat, si%resid_last, si%rtol_last, c0, c1, values, valuesold, d_eta, wv, wv_e, wv_i, cfl_z, rhs_dyn
  if (ab1 == 3) write(u) older
  close(u)
end program fgpu_case_driver
