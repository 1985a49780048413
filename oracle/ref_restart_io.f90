! ref_restart_io.f90 -- TEST INFRASTRUCTURE. A program over the reference's own
! t_mesh, t_tracer and t_dyn modules (MOD_MESH.F90, MOD_TRACER.F90,
! MOD_DYN.F90 with their binary-array helpers, compiled from the reference
! sources into oracle/_ref by oracle/Makefile) that writes and reads the
! derived-type binary restart records the way write_all_bin_restarts /
! read_all_bin_restarts do (io_restart_derivedtype.F90:72-118, 183-230):
! one unformatted sequential record per type, written by the types'
! user-defined derived-type I/O.
!
!   ref_restart_io write <fields.bin> <dir>   fields -> dir/t_mesh.0 t_tracer.0 t_dynamics.0
!   ref_restart_io read  <dir> <fields.bin>   dir/t_*.0 -> fields (same layout)
!
! fields.bin (stream, little-endian; int32 and real64, Fortran array order):
!   int32 N2, F2, E2, nl, kmax, nza, ntr, edge2D_in
!   real  ocean_area, ocean_areawithcav
!   elem2D_nodes(3,F2) edges(2,E2) edge_tri(2,E2) elem_area(F2) edge_dxdy(2,E2)
!   edge_cross_dxdy(4,E2) elem_cos(F2) nod_in_elem2D(kmax,N2) nod_in_elem2D_num(N2)
!   depth(N2) gradient_sca(6,F2) zbar(nl) Z(nl-1) ulevels(F2) ulevels_nod2D(N2)
!   ulevels_nod2D_max(N2) nlevels(F2) nlevels_nod2D(N2) nlevels_nod2D_min(N2)
!   area(nl,N2) areasvol(nl,N2) rowptr(N2+1) colind(nza) values(nza)
!   hnode(nl-1,N2) hnode_new(nl-1,N2) zbar_3d_n(nl,N2) Z_3d_n(nl-1,N2)
!   helem(nl-1,F2) zbar_n_bot(N2) zbar_e_bot(F2) zbar_e_srf(F2)
!   per tracer: int32 ID, i_vert_diff; real ph, pv; char(20) hor, ver, lim;
!               values(nl-1,N2) valuesold(1,nl-1,N2) valuesAB(nl-1,N2)
!   uv(2,nl-1,F2) w(nl,N2) w_e(nl,N2) w_i(nl,N2) cfl_z(nl,N2)
!   int32 use_wsplit, maxiter; real wsplit_maxcfl, soltol
program ref_restart_io
  use MOD_MESH
  use MOD_TRACER
  use MOD_DYN
  implicit none
  character(len=8) :: mode
  character(len=4096) :: a1, a2
  type(t_mesh) :: mesh
  type(t_tracer) :: tracers
  type(t_dyn) :: dyn
  integer :: u
  integer :: N2, F2, E2, nl, kmax, nza, ntr, i, ws, mi
  integer :: sz(8)

  call get_command_argument(1, mode)
  call get_command_argument(2, a1)
  call get_command_argument(3, a2)
  if (trim(mode) == 'write') then
     call read_fields(trim(a1))
     open(newunit=u, file=trim(a2)//'/t_mesh.0', status='replace', form='unformatted')
     write(u) mesh
     close(u)
     open(newunit=u, file=trim(a2)//'/t_tracer.0', status='replace', form='unformatted')
     write(u) tracers
     close(u)
     open(newunit=u, file=trim(a2)//'/t_dynamics.0', status='replace', form='unformatted')
     write(u) dyn
     close(u)
  else if (trim(mode) == 'read') then
     open(newunit=u, file=trim(a1)//'/t_mesh.0', status='old', form='unformatted')
     read(u) mesh
     close(u)
     open(newunit=u, file=trim(a1)//'/t_tracer.0', status='old', form='unformatted')
     read(u) tracers
     close(u)
     open(newunit=u, file=trim(a1)//'/t_dynamics.0', status='old', form='unformatted')
     read(u) dyn
     close(u)
     call write_fields(trim(a2))
  else
     write(*, *) 'usage: ref_restart_io write <fields.bin> <dir> | read <dir> <fields.bin>'
     error stop 2
  end if

contains

  subroutine read_fields(path)
    character(len=*), intent(in) :: path
    integer :: v, iv, itr
    real(WP) :: r2(2)
    open(newunit=v, file=path, access='stream', form='unformatted', status='old', action='read')
    read(v) sz
    N2 = sz(1); F2 = sz(2); E2 = sz(3); nl = sz(4); kmax = sz(5); nza = sz(6); ntr = sz(7)
    mesh%nod2D = N2; mesh%elem2D = F2; mesh%edge2D = E2; mesh%nl = nl; mesh%edge2D_in = sz(8)
    read(v) r2
    mesh%ocean_area = r2(1); mesh%ocean_areawithcav = r2(2)
    allocate(mesh%elem2D_nodes(3, F2), mesh%edges(2, E2), mesh%edge_tri(2, E2), mesh%elem_area(F2), &
         mesh%edge_dxdy(2, E2), mesh%edge_cross_dxdy(4, E2), mesh%elem_cos(F2), mesh%nod_in_elem2D(kmax, N2), &
         mesh%nod_in_elem2D_num(N2), mesh%depth(N2), mesh%gradient_sca(6, F2), mesh%zbar(nl), mesh%Z(nl - 1), &
         mesh%ulevels(F2), mesh%ulevels_nod2D(N2), mesh%ulevels_nod2D_max(N2), mesh%nlevels(F2), &
         mesh%nlevels_nod2D(N2), mesh%nlevels_nod2D_min(N2), mesh%area(nl, N2), mesh%areasvol(nl, N2), &
         mesh%ssh_stiff%rowptr(N2 + 1), mesh%ssh_stiff%colind(nza), mesh%ssh_stiff%values(nza), &
         mesh%hnode(nl - 1, N2), mesh%hnode_new(nl - 1, N2), mesh%zbar_3d_n(nl, N2), mesh%Z_3d_n(nl - 1, N2), &
         mesh%helem(nl - 1, F2), mesh%zbar_n_bot(N2), mesh%zbar_e_bot(F2), mesh%zbar_e_srf(F2))
    read(v) mesh%elem2D_nodes, mesh%edges, mesh%edge_tri, mesh%elem_area, mesh%edge_dxdy, mesh%edge_cross_dxdy, &
         mesh%elem_cos, mesh%nod_in_elem2D, mesh%nod_in_elem2D_num, mesh%depth, mesh%gradient_sca, mesh%zbar, &
         mesh%Z, mesh%ulevels, mesh%ulevels_nod2D, mesh%ulevels_nod2D_max, mesh%nlevels, mesh%nlevels_nod2D, &
         mesh%nlevels_nod2D_min, mesh%area, mesh%areasvol, mesh%ssh_stiff%rowptr, mesh%ssh_stiff%colind, &
         mesh%ssh_stiff%values, mesh%hnode, mesh%hnode_new, mesh%zbar_3d_n, mesh%Z_3d_n, mesh%helem, &
         mesh%zbar_n_bot, mesh%zbar_e_bot, mesh%zbar_e_srf
    mesh%ssh_stiff%dim = N2
    mesh%ssh_stiff%nza = nza
    mesh%ssh_stiff%rowptr_loc = mesh%ssh_stiff%rowptr
    mesh%ssh_stiff%colind_loc = mesh%ssh_stiff%colind
    tracers%num_tracers = ntr
    allocate(tracers%data(ntr))
    do itr = 1, ntr
       associate (d => tracers%data(itr))
         read(v) d%ID, iv, d%tra_adv_ph, d%tra_adv_pv, d%tra_adv_hor, d%tra_adv_ver, d%tra_adv_lim
         d%i_vert_diff = iv == 1
         d%gamma0_tra = 0.0_WP
         d%gamma1_tra = 0.0_WP
         d%gamma2_tra = 0.0_WP
         allocate(d%values(nl - 1, N2), d%valuesold(1, nl - 1, N2), d%valuesAB(nl - 1, N2))
         read(v) d%values, d%valuesold, d%valuesAB
       end associate
    end do
    allocate(dyn%uv(2, nl - 1, F2), dyn%w(nl, N2), dyn%w_e(nl, N2), dyn%w_i(nl, N2), dyn%cfl_z(nl, N2))
    read(v) dyn%uv, dyn%w, dyn%w_e, dyn%w_i, dyn%cfl_z
    read(v) ws, mi, r2
    dyn%use_wsplit = ws == 1
    dyn%solverinfo%maxiter = mi
    dyn%wsplit_maxcfl = r2(1)
    dyn%solverinfo%soltol = r2(2)
    close(v)
  end subroutine read_fields

  subroutine write_fields(path)
    character(len=*), intent(in) :: path
    integer :: v, itr
    N2 = mesh%nod2D; F2 = mesh%elem2D; E2 = mesh%edge2D; nl = mesh%nl
    kmax = size(mesh%nod_in_elem2D, 1); nza = mesh%ssh_stiff%nza; ntr = tracers%num_tracers
    open(newunit=v, file=path, access='stream', form='unformatted', status='replace', action='write')
    write(v) N2, F2, E2, nl, kmax, nza, ntr, mesh%edge2D_in
    write(v) mesh%ocean_area, mesh%ocean_areawithcav
    write(v) mesh%elem2D_nodes, mesh%edges, mesh%edge_tri, mesh%elem_area, mesh%edge_dxdy, mesh%edge_cross_dxdy, &
         mesh%elem_cos, mesh%nod_in_elem2D, mesh%nod_in_elem2D_num, mesh%depth, mesh%gradient_sca, mesh%zbar, &
         mesh%Z, mesh%ulevels, mesh%ulevels_nod2D, mesh%ulevels_nod2D_max, mesh%nlevels, mesh%nlevels_nod2D, &
         mesh%nlevels_nod2D_min, mesh%area, mesh%areasvol, mesh%ssh_stiff%rowptr, mesh%ssh_stiff%colind, &
         mesh%ssh_stiff%values, mesh%hnode, mesh%hnode_new, mesh%zbar_3d_n, mesh%Z_3d_n, mesh%helem, &
         mesh%zbar_n_bot, mesh%zbar_e_bot, mesh%zbar_e_srf
    do itr = 1, ntr
       associate (d => tracers%data(itr))
         write(v) d%ID, merge(1, 0, d%i_vert_diff), d%tra_adv_ph, d%tra_adv_pv, d%tra_adv_hor, d%tra_adv_ver, &
              d%tra_adv_lim
         write(v) d%values, d%valuesold, d%valuesAB
       end associate
    end do
    write(v) dyn%uv, dyn%w, dyn%w_e, dyn%w_i, dyn%cfl_z
    write(v) merge(1, 0, dyn%use_wsplit), dyn%solverinfo%maxiter, dyn%wsplit_maxcfl, dyn%solverinfo%soltol
    close(v)
  end subroutine write_fields

end program ref_restart_io
