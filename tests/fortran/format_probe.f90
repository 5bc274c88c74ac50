! Writes the restart and monitor layouts with a Fortran runtime, so that the
! Python writers can be compared byte for byte.  Sizes and values are fixed:
! nx = 3, ny = 4, Nvel = 2; W(i,j,m) = i + 10 j + 100 m + 0.5,
! h(i,j,k) = i + 10 j + 100 k + 0.25, b = -h.  mass.bin is a closed-mass
! companion of one block: header, M0 = 1.25, ledger, identity(q) = q/2, then
! W, h, b of the cells with a halo of 2 (x periodic, zero beyond the walls).
! values.txt holds one IEEE
! double per line as 16 hex digits; each is written with ES25.16E3,
! 1P,E18.9 and 1P,E12.4 to numbers.txt.  zone.txt is the Tecplot ZONE record.
program format_probe
  implicit none
  integer, parameter :: nx = 3, ny = 4, nv = 2
  real(8) :: W(nx, ny, 4), f(nx, ny, nv, 2), x, zero(4 + 2 * nv)
  integer :: i, j, k, m, u, ios, ii
  integer(8) :: bits
  do m = 1, 4; do j = 1, ny; do i = 1, nx
    W(i, j, m) = i + 10 * j + 100 * m + 0.5d0
  end do; end do; end do
  do k = 1, nv; do j = 1, ny; do i = 1, nx
    f(i, j, k, 1) = i + 10 * j + 100 * k + 0.25d0
    f(i, j, k, 2) = -f(i, j, k, 1)
  end do; end do; end do

  open(newunit=u, file='flow2d.dat', form='unformatted', status='replace')
  write(u) 1
  write(u) nx, ny, 4
  write(u) (((W(i, j, m), i = 1, nx), j = 1, ny), m = 1, 4)
  close(u)

  open(newunit=u, file='vdf2d.dat', form='unformatted', status='replace')
  write(u) 1
  write(u) nx, ny, nv * 2
  write(u) ((((f(i, j, k, m), i = 1, nx), j = 1, ny), k = 1, nv), m = 1, 2)
  close(u)

  ! a stream file: header, then per cell (j outer, i inner) W, h, b
  zero = 0.0d0
  open(newunit=u, file='mass.bin', access='stream', form='unformatted', status='replace')
  write(u) 4, nx, ny, nv, 2, 7, 1.25d0, 0.5d0, -0.5d0, 0.0d0, 1.0d-300, (0.5d0 * k, k = 1, 31)
  do j = -1, ny + 2; do i = -1, nx + 2
    if (j < 1 .or. j > ny) then
      write(u) zero
    else
      ii = modulo(i - 1, nx) + 1
      write(u) W(ii, j, :), f(ii, j, :, 1), f(ii, j, :, 2)
    end if
  end do; end do
  close(u)

  open(newunit=u, file='zone.txt', status='replace')
  write(u, '(a20,I8,a20,I8,a20)') "ZONE  I = ", nx, ", J = ", ny, ", DATAPACKING=BLOCK"
  close(u)

  open(newunit=u, file='values.txt', status='old')
  open(newunit=m, file='numbers.txt', status='replace')
  do
    read(u, '(Z16)', iostat=ios) bits
    if (ios /= 0) exit
    x = transfer(bits, x)
    write(m, '(ES25.16E3,"|",1P,E18.9,"|",E12.4)') x, x, x
  end do
  close(u); close(m)
end program format_probe
