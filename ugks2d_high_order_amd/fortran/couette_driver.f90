! Fortran host of the HighOrderImplicit step: the Solve_UGKS iteration order
! (src/mod_Solve.f90:115-165 of the reference) over the C ABI.
!
!   couette_driver <table.bin> <nx> <ny> <nu> <nv> <omega> <mu_ref> <iterations>
!
! table.bin holds u, v, w (Nvel doubles each, native byte order).  The case is
! the closed Couette flow with diffuse walls at -+3.872983346207417, T = 2,
! Pr = 2/3, CFL 0.1, nCFL 1000, PR(2), one LU-SGS pair and the mass rescale.
! Prints one line per iteration with the residual norms of the accepted state.
program couette_driver
  use iso_c_binding
  use ugks_hip
  implicit none
  type(ugks_params_t) :: p
  type(c_ptr) :: h
  real(c_double), allocatable :: u(:), v(:), w(:), fields(:, :, :)
  real(c_double) :: diag(16), norms(24), walls(16)
  integer(c_int) :: rc, acc
  integer :: nvel, it, niter, unit_no
  character(len=512) :: arg

  call get_command_argument(1, arg)
  p%nx = iarg(2); p%ny = iarg(3); p%nu = iarg(4); p%nv = iarg(5)
  p%omega = rarg(6); p%mu_ref = rarg(7); niter = iarg(8)
  nvel = p%nu * p%nv
  allocate(u(nvel), v(nvel), w(nvel))
  open(newunit=unit_no, file=trim(arg), access="stream", form="unformatted", status="old")
  read(unit_no) u, v, w
  close(unit_no)

  p%ck = 1; p%macro_sweeps = 2; p%micro_sweeps = 1; p%mass_rescale = 1
  p%dx = 1.0d0 / p%nx; p%dy = 1.0d0 / p%ny
  p%cfl = 0.1d0; p%ncfl = 1000.0d0; p%pr = 2.0d0 / 3.0d0
  p%wall_ux = [-3.872983346207417d0, 3.872983346207417d0]
  p%wall_tn = [2.0d0, 2.0d0]
  p%min_alpha = 1.0d-6; p%moment_tol = 1.0d-12; p%residual_tol = 1.0d-8

  rc = ugks_create(p, u, v, w, h)
  call check(rc, "ugks_create")
  rc = ugks_cold_start(h, 1.0d0, 0.0d0, 2.0d0)
  call check(rc, "ugks_cold_start")
  do it = 1, niter
    rc = ugks_implicit_step(h, acc, diag)
    call check(rc, "ugks_implicit_step")
    if (acc /= 1) stop "update rejected"
    ! residuals of the accepted state (its operator is reused by the next update)
    rc = ugks_evolution(h, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, &
                        c_null_ptr)
    call check(rc, "ugks_evolution")
    rc = ugks_norms(h, norms, walls)
    call check(rc, "ugks_norms")
    write(*, "(a,i0,4(1x,es24.16e3),2(1x,es24.16e3),1x,es24.16e3)") "iter ", it, norms(1:4), norms(13:14), diag(3)
  end do
  allocate(fields(10, p%nx, p%ny))
  rc = ugks_observables(h, fields)
  call check(rc, "ugks_observables")
  write(*, "(a,3(1x,es24.16e3))") "fields ", fields(6, 1, 1), fields(7, 1, p%ny / 2), fields(9, 1, p%ny)
  call ugks_destroy(h)

contains

  integer function iarg(k)
    integer, intent(in) :: k
    character(len=64) :: s
    call get_command_argument(k, s)
    read(s, *) iarg
  end function iarg

  real(c_double) function rarg(k)
    integer, intent(in) :: k
    character(len=64) :: s
    call get_command_argument(k, s)
    read(s, *) rarg
  end function rarg

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) then
      write(*, "(a,a,a,i0,a,a)") "failed: ", what, " (", code, ") ", ugks_error_text()
      stop 1
    end if
  end subroutine check

end program couette_driver
