!> Fortran driver of the spectral-division drop-ins, shaped like pqdwheig.f90:
!> write a symmetric matrix A = H diag(lam) H (H a Householder reflector,
!> lam spread over [-1, 2]) as a MatrixMarket file, read it back with
!> prepare_real_symm_MM, take the QDWH polar factor of A - sigma I with
!> PQDWHFACS, run PQDWHSUBIT, solve the leading Rk x Rk block with
!> ELPA_PDSYEVD2 and form V1 VV1 as the reference driver does; then the same
!> through ZoloEig.  Checks Rk, the eigenvalues below sigma against lam and the
!> residual ||A V - V diag(w)||.  Prints PASS or FAIL.
program test_zoloeig
  use zolo_capi
  use mod_prepare_matrix
  implicit none
  integer, parameter :: n = 153
  double precision, allocatable :: a(:, :), as(:, :), q(:, :), z(:, :), v(:, :), work(:), work2(:), tau(:), ev(:)
  double precision :: lam(n), u(n), sigma, norma, werr, res, zero_shift
  integer :: desca(9), descw(9), info, i, j, nnz, rk, nbelow, lwork, lwork2
  logical :: ok
  character(len=128) :: filename
  external :: pqdwhfacs, pqdwhsubit, elpa_pdsyevd2, zoloeig

  allocate (a(n, n), as(n, n), q(n, n), z(n, n), v(n, n), tau(n), ev(n))
  do i = 1, n
    lam(i) = -1d0 + 3d0 * dble(i - 1) / dble(n - 1)
    u(i) = sin(dble(i)) + 0.25d0
  end do
  sigma = 0.5d0 * (lam(60) + lam(61))
  nbelow = 60
  call householder_matrix(n, lam, u, a)

  ! the MatrixMarket file goes where the first argument says (default: the current directory)
  filename = 'zoloeig_test.mtx'
  if (command_argument_count() >= 1) call get_command_argument(1, filename)
  open (unit=11, file=trim(filename), status='replace')
  write (11, '(a)') '%%MatrixMarket matrix coordinate real symmetric'
  write (11, '(3i8)') n, n, n * (n + 1) / 2
  do j = 1, n
    do i = j, n
      write (11, '(2i6,es26.17)') i, j, a(i, j)
    end do
  end do
  close (11)

  desca = (/1, 0, n, n, 32, 32, 0, 0, n/)
  descw = (/1, 0, 2 * n, n, 32, 32, 0, 0, 2 * n/)
  ok = .true.
  a = 0d0
  nnz = n * n
  zero_shift = 1d7
  call prepare_real_symm_mm(filename, n, nnz, 0, 0, 0, 1, 1, desca, a, z, zero_shift)
  ok = ok .and. nnz == n * (n + 1) / 2 .and. zero_shift == 0d0
  as = a
  norma = sqrt(sum(a * a))

  lwork = 2 * n * n
  lwork2 = n * n
  allocate (work(lwork), work2(lwork2))
  call pqdwhfacs('N', n, n, sigma, a, 1, 1, desca, q, 1, 1, desca, work, lwork, descw, tau, work2, lwork2, info, 'N')
  write (*, '(a,i6,a,i3)') 'pqdwhfacs info=', info, ' iterations=', zolo_last_info%iters
  ok = ok .and. info == 0
  call pqdwhsubit(n, a, 1, 1, desca, norma, q, 1, 1, desca, work, lwork, tau, work2, lwork2, rk, info)
  write (*, '(a,i6,a,i4)') 'pqdwhsubit info=', info, ' Rk=', rk
  ok = ok .and. info == 0 .and. rk == nbelow
  if (ok) then
    call elpa_pdsyevd2('L', rk, rk, a, n, n, ev, z, 32, desca, 1, 1, 0, 0, 0, 0, info)
    ok = ok .and. info == 0
    v(:, 1:rk) = matmul(q(:, 1:rk), z(1:rk, 1:rk))
    call check(rk, ev, v, werr, res)
    write (*, '(a,es10.3,a,es10.3)') 'pqdwhsubit + elpa_pdsyevd2: eigenvalue error=', werr, ' residual=', res
    ok = ok .and. werr < 1d-12 .and. res < 1d-12
  end if

  ! ZoloEig on the same polar factor
  a = as
  call pqdwhfacs('N', n, n, sigma, a, 1, 1, desca, q, 1, 1, desca, work, lwork, descw, tau, work2, lwork2, info, 'N')
  ok = ok .and. info == 0
  call zoloeig(n, 32, q, a, 0, 0, work, lwork, work2, lwork2, tau, info)
  rk = zolo_eig_count
  write (*, '(a,i6,a,i4)') 'zoloeig info=', info, ' Rk=', rk
  ok = ok .and. info == 0 .and. rk == nbelow
  if (ok) then
    v(:, 1:rk) = reshape(work(1:n * rk), (/n, rk/))
    call check(rk, tau, v, werr, res)
    write (*, '(a,es10.3,a,es10.3)') 'zoloeig: eigenvalue error=', werr, ' residual=', res
    ok = ok .and. werr < 1d-12 .and. res < 1d-12
  end if
  if (ok) then
    write (*, '(a)') 'PASS'
  else
    write (*, '(a)') 'FAIL'
  end if

contains

  subroutine householder_matrix(m, d, w, out)
    integer, intent(in) :: m
    double precision, intent(in) :: d(m), w(m)
    double precision, intent(out) :: out(m, m)
    double precision :: hh(m, m), t
    integer :: ii
    t = 2d0 / dot_product(w, w)
    hh = 0d0
    do ii = 1, m
      hh(ii, ii) = 1d0
    end do
    do ii = 1, m
      hh(:, ii) = hh(:, ii) - t * w * w(ii)
    end do
    do ii = 1, m
      out(:, ii) = matmul(hh, d * hh(:, ii))
    end do
    out = 0.5d0 * (out + transpose(out))
  end subroutine householder_matrix

  subroutine check(k, w, vv, werr_, res_)
    integer, intent(in) :: k
    double precision, intent(in) :: w(*), vv(n, *)
    double precision, intent(out) :: werr_, res_
    double precision :: r(n, k)
    integer :: jj
    werr_ = maxval(abs(w(1:k) - lam(1:k)))
    r = matmul(as, vv(:, 1:k))
    do jj = 1, k
      r(:, jj) = r(:, jj) - w(jj) * vv(:, jj)
    end do
    res_ = sqrt(sum(r * r)) / norma
  end subroutine check
end program test_zoloeig
