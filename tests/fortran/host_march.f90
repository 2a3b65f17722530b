! host_march.f90 -- a Fortran host driving libucnsgpu.so through the
! ISO_C_BINDING module ucns_gpu, the way UCNS3D's TIME_MARCHING would
! (src/time_p.f90:4026-4580): per step CALCULATE_CFL (ucns_compute_dt), then
! RUNGE_KUTTA3 with that dt (ucns_rk3_steps, one step), T += DT.
!
! The mesh and PRESTORE data come from the repository's preprocessor
! (libucnsprep.so, bound here as a host would bind its own set-up), and the
! host copies them into arrays of its own in Fortran order with 1-based
! cell and face indices (index_base = 1), as it packs IELEM/ILOCAL_RECON3.
!
!   host_march errors                          ABI version and create-time error paths (no GPU needed)
!   host_march march tet|hex n steps out.bin   WENO-4 (tet) or MUSCL (hex) periodic vortex, writes
!                                              ncell (int32), t (real64), U(5, ncell) (real64) to out.bin
program host_march
  use iso_c_binding
  use ucns_gpu
  implicit none

  type, bind(c) :: ucnsp_opts
    integer(c_int32_t) :: shape, nx, ny, nz
    real(c_double) :: lx, ly, lz, jitter
    integer(c_int64_t) :: seed
    integer(c_int32_t) :: iorder, iweno, extf, weight_lsqr, nthreads
    integer(c_int32_t) :: bc(6)
    real(c_double) :: bump(8)
    integer(c_int32_t) :: ees
  end type

  interface
    integer(c_int) function ucnsp_build(opts, case) bind(c, name="ucnsp_build")
      import :: c_int, c_ptr, ucnsp_opts
      type(ucnsp_opts), intent(in) :: opts
      type(c_ptr), intent(out) :: case
    end function
    integer(c_int) function ucnsp_fill(case, precision, gamma, lwci1, mesh, scheme, pre) bind(c, name="ucnsp_fill")
      import :: c_int, c_int32_t, c_ptr, c_double, ucns_mesh_v2, ucns_scheme_v2, ucns_prestore_v1
      type(c_ptr), value :: case
      integer(c_int32_t), value :: precision
      real(c_double), value :: gamma, lwci1
      type(ucns_mesh_v2), intent(out) :: mesh
      type(ucns_scheme_v2), intent(out) :: scheme
      type(ucns_prestore_v1), intent(out) :: pre
    end function
    integer(c_int) function ucnsp_info(case, ints, dbls) bind(c, name="ucnsp_info")
      import :: c_int, c_ptr, c_int64_t, c_double
      type(c_ptr), value :: case
      integer(c_int64_t), intent(out) :: ints(16)
      real(c_double), intent(out) :: dbls(8)
    end function
    integer(c_int) function ucnsp_init_vortex(case, gamma, noise, seed, u) bind(c, name="ucnsp_init_vortex")
      import :: c_int, c_ptr, c_double, c_int64_t
      type(c_ptr), value :: case
      real(c_double), value :: gamma, noise
      integer(c_int64_t), value :: seed
      real(c_double), intent(out) :: u(*)
    end function
    subroutine ucnsp_free(case) bind(c, name="ucnsp_free")
      import :: c_ptr
      type(c_ptr), value :: case
    end subroutine
  end interface

  ! the host's own arrays (Fortran order, 1-based indices)
  integer(c_int32_t), allocatable, target :: ifca(:), ineigh(:, :), ineighn(:, :), bcode(:, :), types(:, :)
  real(c_double), allocatable, target :: angx(:, :), angy(:, :), surf(:, :), vol(:), minedge(:)
  integer(c_int32_t), allocatable, target :: admis(:), full(:), ihexl(:, :, :)
  real(c_double), allocatable, target :: invmat(:, :, :, :), indicator(:, :, :), integ(:, :), vol11(:)
  real(c_double), allocatable, target :: qpoints(:, :, :, :), wq(:), wt(:), u(:, :)

  character(len=64) :: mode
  call get_command_argument(1, mode)
  select case (trim(mode))
  case ("errors")
    call check_errors()
  case ("march")
    call march()
  case default
    error stop "usage: host_march errors | march tet|hex n steps out.bin"
  end select

contains

  subroutine check(rc, what)
    integer(c_int), intent(in) :: rc
    character(len=*), intent(in) :: what
    if (rc /= UCNS_OK) then
      write (*, '(a, a, i0)') what, " failed with code ", rc
      error stop 2
    end if
  end subroutine

  ! a periodic box of the repository's synthetic meshes
  subroutine build_case(shape, n, case)
    character(len=*), intent(in) :: shape
    integer, intent(in) :: n
    type(c_ptr), intent(out) :: case
    type(ucnsp_opts) :: o
    real(c_double), parameter :: two_pi = 6.283185307179586d0
    o%nx = n; o%ny = n; o%nz = n
    o%lx = two_pi; o%ly = two_pi; o%lz = two_pi
    o%weight_lsqr = 0; o%nthreads = 0; o%bc = 0; o%bump = 0.0d0; o%ees = 0
    if (shape == "tet") then       ! 6 Kuhn tetrahedra per hex, jittered, WENO-4
      o%shape = 2; o%jitter = 0.1d0; o%seed = 2; o%iorder = 3; o%iweno = 1; o%extf = 2
    else                           ! hexahedra, MUSCL + Barth
      o%shape = 1; o%jitter = 0.0d0; o%seed = 1; o%iorder = 1; o%iweno = -1; o%extf = 3
    end if
    call check(ucnsp_build(o, case), "ucnsp_build")
  end subroutine

  ! copy the preprocessor's C arrays into the host's Fortran arrays, indices +1
  subroutine pack_host(case, mesh, scheme, pre, nc)
    type(c_ptr), intent(in) :: case
    type(ucns_mesh_v2), intent(out) :: mesh
    type(ucns_scheme_v2), intent(out) :: scheme
    type(ucns_prestore_v1), intent(out) :: pre
    integer, intent(out) :: nc
    type(ucns_mesh_v2) :: cm
    type(ucns_prestore_v1) :: cp
    integer(c_int64_t) :: ints(16)
    real(c_double) :: dbls(8)
    integer :: nf, ideg, nnb, imax, ts, nqm
    integer(c_int32_t), pointer :: pi1(:)
    real(c_double), pointer :: pd1(:)
    call check(ucnsp_fill(case, UCNS_FP64, 1.4d0, 10000.0d0, cm, scheme, cp), "ucnsp_fill")
    call check(ucnsp_info(case, ints, dbls), "ucnsp_info")
    nc = int(ints(1)); nf = int(ints(3)); ideg = int(ints(5)); nnb = int(ints(6))
    imax = int(ints(7)); ts = int(ints(8)); nqm = int(ints(13))
    allocate(ifca(nc), ineigh(nf, nc), ineighn(nf, nc), bcode(nf, nc), types(nf, nc))
    allocate(angx(nf, nc), angy(nf, nc), surf(nf, nc), vol(nc), minedge(nc))
    allocate(admis(nc), full(nc), ihexl(nnb, ts, nc), invmat(ideg, imax, ts, nc), indicator(ideg, ideg, nc))
    allocate(integ(ideg, nc), vol11(nc), qpoints(3, nqm, nf, nc))
    call c_f_pointer(cm%ifca, pi1, [nc]);              ifca = pi1
    call c_f_pointer(cm%ineigh, pi1, [nf * nc]);       ineigh = reshape(pi1, [nf, nc]) + 1
    call c_f_pointer(cm%ineighn, pi1, [nf * nc]);      ineighn = reshape(pi1, [nf, nc]) + 1
    call c_f_pointer(cm%bcode, pi1, [nf * nc]);        bcode = reshape(pi1, [nf, nc])
    call c_f_pointer(cm%types_faces, pi1, [nf * nc]);  types = reshape(pi1, [nf, nc])
    call c_f_pointer(cm%faceanglex, pd1, [nf * nc]);   angx = reshape(pd1, [nf, nc])
    call c_f_pointer(cm%faceangley, pd1, [nf * nc]);   angy = reshape(pd1, [nf, nc])
    call c_f_pointer(cm%surf, pd1, [nf * nc]);         surf = reshape(pd1, [nf, nc])
    call c_f_pointer(cm%totvolume, pd1, [nc]);         vol = pd1
    call c_f_pointer(cm%minedge, pd1, [nc]);           minedge = pd1
    call c_f_pointer(cp%admis, pi1, [nc]);             admis = pi1
    call c_f_pointer(cp%full, pi1, [nc]);              full = pi1
    call c_f_pointer(cp%ihexl, pi1, [nnb * ts * nc]);  ihexl = reshape(pi1, [nnb, ts, nc]) + 1
    call c_f_pointer(cp%invmat, pd1, [ideg * imax * ts * nc]);  invmat = reshape(pd1, [ideg, imax, ts, nc])
    if (c_associated(cp%indicator)) then
      call c_f_pointer(cp%indicator, pd1, [ideg * ideg * nc]);  indicator = reshape(pd1, [ideg, ideg, nc])
    else
      indicator = 0.0d0
    end if
    call c_f_pointer(cp%integ_basis, pd1, [ideg * nc]);         integ = reshape(pd1, [ideg, nc])
    call c_f_pointer(cp%volume11, pd1, [nc]);                   vol11 = pd1
    call c_f_pointer(cp%qpoints, pd1, [3 * nqm * nf * nc]);     qpoints = reshape(pd1, [3, nqm, nf, nc])
    allocate(wq(scheme%qp_quad), wt(scheme%qp_triangle))
    call c_f_pointer(scheme%weights_quad, pd1, [scheme%qp_quad]);      wq = pd1
    call c_f_pointer(scheme%weights_tri, pd1, [scheme%qp_triangle]);   wt = pd1
    mesh%abi_version = UCNS_ABI_VERSION
    mesh%ncell = nc; mesh%index_base = 1; mesh%nface_max = nf
    mesh%ifca = c_loc(ifca); mesh%ineigh = c_loc(ineigh); mesh%ineighn = c_loc(ineighn)
    mesh%bcode = c_loc(bcode); mesh%types_faces = c_loc(types)
    mesh%faceanglex = c_loc(angx); mesh%faceangley = c_loc(angy); mesh%surf = c_loc(surf)
    mesh%totvolume = c_loc(vol); mesh%minedge = c_loc(minedge); mesh%centroid = c_null_ptr
    scheme%weights_quad = c_loc(wq); scheme%weights_tri = c_loc(wt)
    pre%admis = c_loc(admis); pre%full = c_loc(full); pre%ihexl = c_loc(ihexl); pre%invmat = c_loc(invmat)
    pre%indicator = c_loc(indicator); pre%integ_basis = c_loc(integ); pre%volume11 = c_loc(vol11)
    pre%qpoints = c_loc(qpoints)
  end subroutine

  subroutine expect_create_error(mesh, scheme, want, needle)
    type(ucns_mesh_v2), intent(in) :: mesh
    type(ucns_scheme_v2), intent(in) :: scheme
    integer(c_int), intent(in) :: want
    character(len=*), intent(in) :: needle
    type(c_ptr) :: ctx
    integer(c_int) :: rc
    character(len=:), allocatable :: msg
    rc = ucns_create(mesh, scheme, ctx)
    msg = ucns_error_message(ctx)
    call ucns_destroy(ctx)
    write (*, '(a, i0, 3a)') "  create: code ", rc, " (", msg, ")"
    if (rc /= want .or. index(msg, needle) == 0) then
      write (*, '(3a, i0)') "expected '", needle, "' with code ", want
      error stop 3
    end if
  end subroutine

  subroutine check_errors()
    type(c_ptr) :: case
    type(ucns_mesh_v2) :: mesh, m2
    type(ucns_scheme_v2) :: scheme, s2
    type(ucns_prestore_v1) :: pre
    integer :: nc
    real(c_double) :: rhs(5)
    if (ucns_library_abi() /= UCNS_ABI_VERSION) error stop "ABI version of the library differs from the module's"
    write (*, '(a, i0)') "  library ABI version ", ucns_library_abi()
    call build_case("hex", 3, case)
    call pack_host(case, mesh, scheme, pre, nc)
    m2 = mesh; m2%abi_version = 99
    call expect_create_error(m2, scheme, UCNS_EBADARG, "abi_version")
    m2 = mesh; m2%surf = c_null_ptr
    call expect_create_error(m2, scheme, UCNS_EBADARG, "null mesh array")
    s2 = scheme; s2%ees = 3
    call expect_create_error(mesh, s2, UCNS_EUNSUPPORTED, "ees must be 0")
    s2 = scheme; s2%cell_order = 7
    call expect_create_error(mesh, s2, UCNS_EBADARG, "cell_order")
    s2 = scheme; s2%cell_order = UCNS_ORDER_MORTON
    call expect_create_error(mesh, s2, UCNS_EBADARG, "centroids")
    m2 = mesh; ineigh(1, 1) = nc + 5
    call expect_create_error(m2, scheme, UCNS_EBADARG, "face neighbour out of range")
    if (ucns_residual(c_null_ptr, rhs, UCNS_LAYOUT_AOS) /= UCNS_EBADARG) error stop "null context accepted"
    if (ucns_error_message(c_null_ptr) /= "null context") error stop "null-context message"
    call ucnsp_free(case)
    print '(a)', "FORTRAN_ERRORS_OK"
  end subroutine

  subroutine march()
    character(len=16) :: shape
    character(len=32) :: arg
    character(len=512) :: out
    integer :: n, nsteps, it, nc, unit
    type(c_ptr) :: case, ctx
    type(ucns_mesh_v2) :: mesh
    type(ucns_scheme_v2) :: scheme
    type(ucns_prestore_v1) :: pre
    real(c_double) :: t, dt
    real(c_double), parameter :: cfl = 0.5d0
    call get_command_argument(2, shape)
    call get_command_argument(3, arg); read (arg, *) n
    call get_command_argument(4, arg); read (arg, *) nsteps
    call get_command_argument(5, out)
    call build_case(trim(shape), n, case)
    call pack_host(case, mesh, scheme, pre, nc)
    allocate(u(5, nc))
    call check(ucnsp_init_vortex(case, 1.4d0, 0.0d0, 1_c_int64_t, u), "ucnsp_init_vortex")
    call check(ucns_create(mesh, scheme, ctx), "ucns_create")
    call check(ucns_upload_prestore(ctx, pre), "ucns_upload_prestore")
    call check(ucns_set_state(ctx, u, UCNS_LAYOUT_AOS), "ucns_set_state")
    t = 0.0d0
    do it = 1, nsteps                   ! TIME_MARCHING
      call check(ucns_compute_dt(ctx, cfl, dt), "ucns_compute_dt")            ! CALCULATE_CFL
      call check(ucns_rk3_steps(ctx, 1_c_int32_t, cfl, dt, t), "ucns_rk3_steps")  ! RUNGE_KUTTA3, T = T + DT
    end do
    call check(ucns_get_state(ctx, u, UCNS_LAYOUT_AOS), "ucns_get_state")
    call ucns_destroy(ctx)
    call ucnsp_free(case)
    open (newunit=unit, file=trim(out), access="stream", form="unformatted", status="replace")
    write (unit) int(nc, c_int32_t), t, u
    close (unit)
    write (*, '(a, i0, a, i0, a, es23.16)') "FORTRAN_MARCH_OK ncell ", nc, " steps ", nsteps, " t ", t
  end subroutine

end program host_march
