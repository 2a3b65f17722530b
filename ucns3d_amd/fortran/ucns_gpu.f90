! ucns_gpu.f90 -- ISO_C_BINDING interface of libucnsgpu.so (include/ucns.h, ABI 2)
! for a Fortran host.  The host keeps its mesh, UCNS3D.DAT switches and
! set-up, packs the IELEM / ILOCAL_RECON3 fields it needs into plain arrays
! (index_base = 1 for Fortran indices) and calls these in place of
! CALL_FLUX_SUBROUTINES_3D (src/time_p.f90:2253-2497), the stage loops of
! RUNGE_KUTTA3 (src/time_p.f90:885-1027) and the dt loop of CALCULATE_CFL
! (src/time_p.f90:27-191).  Every function returns UCNS_OK (0) or a negative
! code; ucns_error_message gives the text.  C row-major [a][b][c] is the
! Fortran array x(c, b, a).
module ucns_gpu
  use iso_c_binding
  implicit none
  private :: c_strlen

  integer(c_int32_t), parameter :: UCNS_ABI_VERSION = 2
  integer(c_int), parameter :: UCNS_OK = 0, UCNS_EBADARG = -1, UCNS_EUNSUPPORTED = -2, UCNS_ENOMEM = -3
  integer(c_int), parameter :: UCNS_EHIP = -4, UCNS_ENONFINITE = -5, UCNS_ESTATE = -6
  integer(c_int32_t), parameter :: UCNS_FP64 = 0, UCNS_MIXED = 1
  integer(c_int32_t), parameter :: UCNS_LAYOUT_AOS = 0, UCNS_LAYOUT_SOA = 1
  integer(c_int32_t), parameter :: UCNS_ORDER_CALLER = 0, UCNS_ORDER_RCM = 1, UCNS_ORDER_MORTON = 2
  integer(c_int32_t), parameter :: UCNS_KERNELS_AUTO = 0, UCNS_KERNELS_REGISTER = 1

  type, bind(c) :: ucns_mesh_v2
    integer(c_int32_t) :: abi_version = UCNS_ABI_VERSION, ncell = 0, index_base = 1, nface_max = 0
    type(c_ptr) :: ifca = c_null_ptr, ineigh = c_null_ptr, ineighn = c_null_ptr, bcode = c_null_ptr
    type(c_ptr) :: types_faces = c_null_ptr, faceanglex = c_null_ptr, faceangley = c_null_ptr
    type(c_ptr) :: surf = c_null_ptr, totvolume = c_null_ptr, minedge = c_null_ptr, centroid = c_null_ptr
  end type

  type, bind(c) :: ucns_scheme_v2
    integer(c_int32_t) :: abi_version = UCNS_ABI_VERSION, precision = UCNS_FP64, iweno = 1, iorder = 0
    integer(c_int32_t) :: idegfree = 0, inumneighbours = 0, typesten = 0, ees = 0, wenoz = 0, wenwrt = 1
    integer(c_int32_t) :: limiter = 1, riemann = 1, qp_quad = 0, qp_triangle = 0
    real(c_double) :: gamma = 1.4d0, lwci1 = 1.0d0, weno_eps = 1.0d-11
    type(c_ptr) :: weights_quad = c_null_ptr, weights_tri = c_null_ptr
    integer(c_int32_t) :: cell_order = UCNS_ORDER_CALLER, kernels = UCNS_KERNELS_AUTO
    integer(c_int32_t) :: adj_keep_stages = 0, reserved = 0
    real(c_double) :: adj_keep_gib = 0.0d0
  end type

  type, bind(c) :: ucns_prestore_v1
    type(c_ptr) :: admis = c_null_ptr, full = c_null_ptr, ihexl = c_null_ptr, invmat = c_null_ptr
    type(c_ptr) :: indicator = c_null_ptr, integ_basis = c_null_ptr, volume11 = c_null_ptr, qpoints = c_null_ptr
  end type

  type, bind(c) :: ucns_boundary_v1
    integer(c_int32_t) :: abi_version = UCNS_ABI_VERSION, boundtype = 0
    real(c_double) :: inflow(5) = 0.0d0, outflow(5) = 0.0d0
  end type

  ! CWENO data (ees = 5): the sectors' low-order pseudo-inverses, indicators
  ! and basis integrals, c_loc() of the host's packed arrays
  type, bind(c) :: ucns_cweno_v1
    integer(c_int32_t) :: abi_version = UCNS_ABI_VERSION
    integer(c_int32_t) :: idegfree2 = 3, inumneighbours2 = 12, reserved = 0
    type(c_ptr) :: invmat_c = c_null_ptr, indicator_c = c_null_ptr, integ_basis_c = c_null_ptr
  end type

  ! cotangents of the mesh/PRESTORE data (ucns_rk3_adjoint_geometry): c_loc()
  ! of the host's arrays in the layouts of ucns_prestore_v1 / ucns_mesh_v2, or
  ! c_null_ptr for one not wanted
  type, bind(c) :: ucns_geometry_bar_v1
    integer(c_int32_t) :: abi_version = UCNS_ABI_VERSION
    integer(c_int32_t) :: reserved = 0
    type(c_ptr) :: invmat = c_null_ptr, indicator = c_null_ptr, integ_basis = c_null_ptr, volume11 = c_null_ptr
    type(c_ptr) :: qpoints = c_null_ptr, surf = c_null_ptr, normal = c_null_ptr, totvolume = c_null_ptr
  end type

  type, bind(c) :: ucns_stats_v1
    integer(c_int64_t) :: n_recon, n_flux, n_update, n_dt
    real(c_double) :: ms_recon, ms_flux, ms_update, ms_dt, bytes_recon_per_cell
    integer(c_int64_t) :: n_reduced_last, n_firstorder_last, n_recon_vjp, n_flux_vjp
    real(c_double) :: ms_recon_vjp, ms_flux_vjp
    integer(c_int64_t) :: n_recon_adj
    real(c_double) :: ms_recon_adj
  end type

  interface
    ! ucns_abi_version() of the C ABI (Fortran names are case-blind, so it
    ! cannot share its name with the UCNS_ABI_VERSION parameter)
    integer(c_int) function ucns_library_abi() bind(c, name="ucns_abi_version")
      import :: c_int
    end function
    integer(c_int) function ucns_create(mesh, scheme, ctx) bind(c, name="ucns_create")
      import :: ucns_mesh_v2, ucns_scheme_v2, c_int, c_ptr
      type(ucns_mesh_v2), intent(in) :: mesh
      type(ucns_scheme_v2), intent(in) :: scheme
      type(c_ptr), intent(out) :: ctx
    end function
    integer(c_int) function ucns_upload_prestore(ctx, pre) bind(c, name="ucns_upload_prestore")
      import :: ucns_prestore_v1, c_int, c_ptr
      type(c_ptr), value :: ctx
      type(ucns_prestore_v1), intent(in) :: pre
    end function
    integer(c_int) function ucns_upload_cweno(ctx, cw) bind(c, name="ucns_upload_cweno")
      import :: c_int, c_ptr, ucns_cweno_v1
      type(c_ptr), value :: ctx
      type(ucns_cweno_v1), intent(in) :: cw
    end function
    integer(c_int) function ucns_set_boundary(ctx, bc) bind(c, name="ucns_set_boundary")
      import :: ucns_boundary_v1, c_int, c_ptr
      type(c_ptr), value :: ctx
      type(ucns_boundary_v1), intent(in) :: bc
    end function
    ! per-face INFLOW()/OUTFLOW() states: c_loc of real(c_double) arrays (5, nface_max, ncell), or c_null_ptr
    integer(c_int) function ucns_set_boundary_faces(ctx, inflow, outflow) bind(c, name="ucns_set_boundary_faces")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, inflow, outflow
    end function
    integer(c_int) function ucns_set_state(ctx, u, layout) bind(c, name="ucns_set_state")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: u(*)
      integer(c_int32_t), value :: layout
    end function
    integer(c_int) function ucns_get_state(ctx, u, layout) bind(c, name="ucns_get_state")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: u(*)
      integer(c_int32_t), value :: layout
    end function
    integer(c_int) function ucns_set_state_device(ctx, u_dev, layout) bind(c, name="ucns_set_state_device")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, u_dev
      integer(c_int32_t), value :: layout
    end function
    integer(c_int) function ucns_get_state_device(ctx, u_dev, layout) bind(c, name="ucns_get_state_device")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, u_dev
      integer(c_int32_t), value :: layout
    end function
    integer(c_int) function ucns_compute_dt(ctx, cfl, dt) bind(c, name="ucns_compute_dt")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), value :: cfl
      real(c_double), intent(out) :: dt
    end function
    integer(c_int) function ucns_residual(ctx, rhs, layout) bind(c, name="ucns_residual")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: rhs(*)
      integer(c_int32_t), value :: layout
    end function
    integer(c_int) function ucns_residual_device(ctx, rhs_dev, layout) bind(c, name="ucns_residual_device")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, rhs_dev
      integer(c_int32_t), value :: layout
    end function
    integer(c_int) function ucns_residual_noalloc(ctx, rhs_dev, layout) bind(c, name="ucns_residual_noalloc")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx, rhs_dev
      integer(c_int32_t), value :: layout
    end function
    integer(c_int) function ucns_reconstruct(ctx, uleft) bind(c, name="ucns_reconstruct")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: uleft(*)
    end function
    integer(c_int) function ucns_rk3_steps(ctx, nsteps, cfl, dt_fixed, t) bind(c, name="ucns_rk3_steps")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
      real(c_double), value :: cfl, dt_fixed
      real(c_double), intent(inout) :: t
    end function
    integer(c_int) function ucns_residual_vjp(ctx, w, ubar, layout, bcbar) bind(c, name="ucns_residual_vjp")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(in) :: w(*)
      real(c_double), intent(out) :: ubar(*)
      integer(c_int32_t), value :: layout
      real(c_double), intent(out) :: bcbar(10)
    end function
    integer(c_int) function ucns_residual_vjp_device(ctx, w_dev, ubar_dev, bcbar_dev) &
        bind(c, name="ucns_residual_vjp_device")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, w_dev, ubar_dev, bcbar_dev
    end function
    integer(c_int) function ucns_rk3_adjoint(ctx, nsteps, cfl, dt_fixed, unbar, u0bar, layout, bcbar, t) &
        bind(c, name="ucns_rk3_adjoint")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
      real(c_double), value :: cfl, dt_fixed
      real(c_double), intent(in) :: unbar(*)
      real(c_double), intent(out) :: u0bar(*)
      integer(c_int32_t), value :: layout
      real(c_double), intent(out) :: bcbar(10)
      real(c_double), intent(inout) :: t
    end function
    integer(c_int) function ucns_rk3_adjoint_geometry(ctx, nsteps, dt_fixed, unbar, u0bar, layout, bcbar, gbar, t) &
        bind(c, name="ucns_rk3_adjoint_geometry")
      import :: c_int, c_int32_t, c_ptr, c_double, ucns_geometry_bar_v1
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
      real(c_double), value :: dt_fixed
      real(c_double), intent(in) :: unbar(*)
      real(c_double), intent(out) :: u0bar(*)
      integer(c_int32_t), value :: layout
      real(c_double), intent(out) :: bcbar(10)
      type(ucns_geometry_bar_v1), intent(in) :: gbar
      real(c_double), intent(inout) :: t
    end function
    integer(c_int) function ucns_rk3_adjoint_device(ctx, nsteps, cfl, dt_fixed, unbar_dev, u0bar_dev, bcbar_dev, t) &
        bind(c, name="ucns_rk3_adjoint_device")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
      real(c_double), value :: cfl, dt_fixed
      type(c_ptr), value :: unbar_dev, u0bar_dev, bcbar_dev
      real(c_double), intent(inout) :: t
    end function
    integer(c_int) function ucns_stats(ctx, st) bind(c, name="ucns_stats")
      import :: c_int, c_ptr, ucns_stats_v1
      type(c_ptr), value :: ctx
      type(ucns_stats_v1), intent(out) :: st
    end function
    integer(c_int) function ucns_reset_stats(ctx) bind(c, name="ucns_reset_stats")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function ucns_debug_fail_alloc(ctx, nth) bind(c, name="ucns_debug_fail_alloc")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nth
    end function
    type(c_ptr) function ucns_last_error(ctx) bind(c, name="ucns_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function
    subroutine ucns_destroy(ctx) bind(c, name="ucns_destroy")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end subroutine
    integer(c_size_t) function c_strlen(s) bind(c, name="strlen")
      import :: c_size_t, c_ptr
      type(c_ptr), value :: s
    end function
  end interface

contains

  ! the message of the context's last failed call, as a Fortran string
  function ucns_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=:), allocatable :: msg
    type(c_ptr) :: p
    character(kind=c_char), pointer :: chars(:)
    integer :: k, n
    p = ucns_last_error(ctx)
    if (.not. c_associated(p)) then
      msg = ""
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate(character(len=n) :: msg)
    do k = 1, n
      msg(k:k) = chars(k)
    end do
  end function

end module ucns_gpu
