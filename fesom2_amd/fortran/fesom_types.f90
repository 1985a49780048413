! fesom_types.f90 -- stand-ins for the FESOM2 modules whose derived types and
! module variables the drop-in entry points (fesom_gpu_dropin.f90) take:
!   o_PARAM, MOD_PARSUP, MOD_PARTIT, MOD_MESH, MOD_TRACER, MOD_DYN, MOD_ICE,
!   o_ARRAYS, g_forcing_arrays, g_config.
! Each declares, under the reference's names and with its kinds and shapes,
! the components and variables that the tracer transport, the implicit
! vertical diffusion and the SSH solve read or write:
!   MOD_MESH.F90:10-126, MOD_TRACER.F90:9-82, MOD_DYN.F90:13-170,
!   MOD_PARTIT.F90 (t_partit sizes), MOD_ICE.F90 (ice%data(1)%values),
!   oce_modules.F90:7-330 (o_PARAM, o_ARRAYS), gen_modules_forcing.F90:72-108,
!   gen_modules_config.F90:55-188.
! A FESOM2 build compiles the drop-in against its own modules of these names
! instead; these exist so that the drop-in and its case driver build and run
! on their own.
module o_PARAM
  use, intrinsic :: iso_fortran_env, only: real64
  implicit none
  integer, parameter :: WP = real64
  logical :: Redi = .false.
end module o_PARAM

module MOD_PARSUP
end module MOD_PARSUP

module MOD_PARTIT
  implicit none
  ! one partition: the halo sizes (eDim_*) stay zero and npes is 1
  type t_partit
     integer :: MPI_COMM_FESOM = 0
     integer :: mype = 0, npes = 1
     integer :: myDim_nod2D = 0, eDim_nod2D = 0
     integer :: myDim_elem2D = 0, eDim_elem2D = 0, eXDim_elem2D = 0
     integer :: myDim_edge2D = 0, eDim_edge2D = 0
  end type t_partit
end module MOD_PARTIT

module MOD_MESH
  use o_PARAM
  use, intrinsic :: iso_fortran_env, only: int32
  implicit none
  type sparse_matrix
     integer :: nza = 0
     integer :: dim = 0
     real(kind=WP), allocatable, dimension(:) :: values
     integer(int32), allocatable, dimension(:) :: colind
     integer(int32), allocatable, dimension(:) :: rowptr
     integer(int32), allocatable, dimension(:) :: colind_loc
     integer(int32), allocatable, dimension(:) :: rowptr_loc
     real(kind=WP), allocatable, dimension(:) :: pr_values
  end type sparse_matrix

  type t_mesh
     integer :: nod2D = 0
     real(kind=WP), allocatable, dimension(:,:) :: coord_nod2D
     integer :: edge2D = 0
     integer :: edge2D_in = 0
     integer :: elem2D = 0
     integer, allocatable, dimension(:,:) :: elem2D_nodes
     integer, allocatable, dimension(:,:) :: edges
     integer, allocatable, dimension(:,:) :: edge_tri
     real(kind=WP), allocatable, dimension(:) :: elem_area
     real(kind=WP), allocatable, dimension(:,:) :: edge_dxdy
     real(kind=WP), allocatable, dimension(:,:) :: edge_cross_dxdy
     real(kind=WP), allocatable, dimension(:) :: elem_cos
     integer, allocatable, dimension(:,:) :: nod_in_elem2D
     integer, allocatable, dimension(:) :: nod_in_elem2D_num
     real(kind=WP), allocatable, dimension(:,:) :: gradient_sca
     integer :: nl = 0
     integer, allocatable, dimension(:) :: ulevels
     integer, allocatable, dimension(:) :: ulevels_nod2D
     integer, allocatable, dimension(:) :: ulevels_nod2D_max
     integer, allocatable, dimension(:) :: nlevels
     integer, allocatable, dimension(:) :: nlevels_nod2D
     integer, allocatable, dimension(:) :: nlevels_nod2D_min
     real(kind=WP), allocatable, dimension(:,:) :: area
     real(kind=WP), allocatable, dimension(:,:) :: areasvol
     type(sparse_matrix) :: ssh_stiff
     real(kind=WP), allocatable, dimension(:,:) :: hnode
     real(kind=WP), allocatable, dimension(:,:) :: hnode_new
     real(kind=WP), allocatable, dimension(:,:) :: zbar_3d_n
     real(kind=WP), allocatable, dimension(:,:) :: Z_3d_n
     real(kind=WP), allocatable, dimension(:,:) :: helem
     real(kind=WP), allocatable, dimension(:) :: zbar_n_bot
     real(kind=WP), allocatable, dimension(:) :: zbar_e_bot
     real(kind=WP), allocatable, dimension(:) :: zbar_e_srf
  end type t_mesh
end module MOD_MESH

module MOD_TRACER
  use o_PARAM
  implicit none
  type t_tracer_data
     real(kind=WP), allocatable, dimension(:,:) :: values
     real(kind=WP), allocatable, dimension(:,:) :: valuesAB
     real(kind=WP), allocatable, dimension(:,:,:) :: valuesold
     logical :: i_vert_diff = .false.
     character(20) :: tra_adv_hor, tra_adv_ver, tra_adv_lim
     real(kind=WP) :: tra_adv_ph = 1.0_WP
     real(kind=WP) :: tra_adv_pv = 1.0_WP
     integer :: AB_order = 2
     integer :: ID
     logical :: ltra_diag = .true.
  end type t_tracer_data

  type t_tracer_work
     real(kind=WP), allocatable :: del_ttf(:,:)
     real(kind=WP), allocatable :: del_ttf_advhoriz(:,:), del_ttf_advvert(:,:)
     real(kind=WP), allocatable, dimension(:,:) :: fct_LO
     real(kind=WP), allocatable, dimension(:,:) :: adv_flux_hor
     real(kind=WP), allocatable, dimension(:,:) :: adv_flux_ver
  end type t_tracer_work

  type t_tracer
     integer :: num_tracers = 2
     type(t_tracer_data), allocatable :: data(:)
     type(t_tracer_work) :: work
  end type t_tracer
end module MOD_TRACER

module MOD_DYN
  use o_PARAM
  implicit none
  type t_solverinfo
     integer :: ident = 1
     integer :: maxiter = 2000
     integer :: restart = 15
     integer :: fillin = 3
     integer :: lutype = 2
     real(kind=WP) :: droptol = 1.e-8_WP
     real(kind=WP) :: soltol = 1e-5_WP
     integer :: precond_variant = 1
     real(kind=WP), allocatable :: rr(:), zz(:), pp(:), App(:)
     integer :: iters_last = 0
     integer :: iters_max = 0
     integer :: iters_sum = 0
     integer :: nsolves = 0
     integer :: nonconv = 0
     integer :: nbreakdown = 0
     integer :: nnegrz = 0
     real(kind=WP) :: resid_last = 0.0_WP
     real(kind=WP) :: rtol_last = 0.0_WP
  end type t_solverinfo

  type t_dyn
     real(kind=WP), allocatable, dimension(:,:,:) :: uv, uv_rhs
     real(kind=WP), allocatable, dimension(:,:) :: w, w_e, w_i, cfl_z
     real(kind=WP), allocatable, dimension(:) :: eta_n, d_eta, ssh_rhs, ssh_rhs_old
     type(t_solverinfo) :: solverinfo
     logical :: use_wsplit = .false.
     real(kind=WP) :: wsplit_maxcfl = 1.0_WP
  end type t_dyn
end module MOD_DYN

module MOD_ICE
  use o_PARAM
  implicit none
  type t_ice_data
     real(kind=WP), allocatable, dimension(:) :: values
  end type t_ice_data
  type t_ice
     type(t_ice_data), allocatable :: data(:)
  end type t_ice
end module MOD_ICE

module o_ARRAYS
  use o_PARAM
  implicit none
  real(kind=WP), allocatable :: heat_flux(:)
  real(kind=WP), allocatable :: water_flux(:)
  real(kind=WP), allocatable :: hosing_flux(:)
  real(kind=WP), allocatable :: hosing_flux3D(:,:)
  real(kind=WP), allocatable :: virtual_salt(:), relax_salt(:)
  real(kind=WP), allocatable, dimension(:,:) :: Kv
  real(kind=WP), allocatable, dimension(:,:,:) :: slope_tapered
  real(kind=WP), allocatable :: Ki(:,:)
  real(kind=WP) :: is_nonlinfs = 0.0_WP
  real(kind=WP), allocatable :: wiso_flux_oce(:,:)
end module o_ARRAYS

module g_forcing_arrays
  use o_PARAM
  implicit none
  real(kind=WP), allocatable, dimension(:) :: prec_rain
  real(kind=WP), allocatable, dimension(:) :: real_salt_flux
end module g_forcing_arrays

module g_config
  use o_PARAM
  implicit none
  character(20) :: which_ALE = 'linfs'
  ! radians once the model setup has scaled it (gen_model_setup.F90:97)
  real(kind=WP) :: cyclic_length = 0.0_WP
  logical :: use_hosing = .false.
  character(10) :: hosing_mode = 'surf'
  real(kind=WP) :: dt
end module g_config
