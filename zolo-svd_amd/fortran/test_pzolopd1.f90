!> Fortran driver of the drop-in entry points, shaped like test_zolodrv1.f90 and
!> pqdwh_fall.f90: build A = Q1 diag(s) Q2^T (Q1, Q2 products of seeded
!> Householder reflectors), set 1x1 descriptors by hand, get alpha/beta, call
!> PZOLOPD1 (FP64 and BF16x3), PQDWHFACS with its workspace query, and check
!> orthogonality and the backward error.  Prints PASS or FAIL.
program test_pzolopd1
  use zolo_capi
  use auxil_zolo
  implicit none
  integer, parameter :: n = 200
  double precision, allocatable :: a(:, :), q(:, :), h(:, :), work(:), work2(:), tau(:)
  double precision :: s(n), alpha, beta, coeff(24), rel, orth
  integer :: desca(9), descq(9), descw(9), info, rk, i, lwork, lwork2
  logical :: ok
  external :: pzolopd1, pqdwhfacs

  allocate (a(n, n), q(n, n), h(n, n), tau(n))
  do i = 1, n
    s(i) = 10d0 ** (-8d0 * dble(i - 1) / dble(n - 1))
  end do
  call make_matrix(n, s, a)
  desca = (/1, 0, n, n, 64, 64, 0, 0, n/)
  descq = desca
  descw = (/1, 0, 2 * n, n, 64, 64, 0, 0, 2 * n/)
  ok = .true.

  call zolo_bounds(0, n, a, n, alpha, beta, info)
  rk = choosem(1d0 / beta)
  call compute_coeff(1d0 / beta, epsilon(1d0), rk, coeff)
  write (*, '(a,es12.4,a,es12.4,a,i2,a,es12.4)') 'alpha=', alpha, ' beta=', beta, ' r=', rk, ' c1=', coeff(1)

  do i = 0, 2, 2
    call zolo_set_precision(i)
    q = 0d0
    call pzolopd1('N', rk, n, n, alpha, beta, a, 1, 1, desca, q, 1, 1, descq, info, 'N')
    call checks(q, rel, orth)
    write (*, '(a,i1,a,i6,a,16a1,a,es10.3,a,es10.3)') 'pzolopd1 prec=', i, ' info=', info, ' schedule=', &
      zolo_last_info%schedule, ' bwd=', rel, ' orth=', orth
    ok = ok .and. info == 0 .and. rel < 1d-13 .and. orth < 1d-13
  end do
  call zolo_set_precision(0)

  ! argument checks: RK out of range, bad descriptor LLD
  call pzolopd1('N', 9, n, n, alpha, beta, a, 1, 1, desca, q, 1, 1, descq, info, 'N')
  ok = ok .and. info == -2
  descq(9) = n - 1
  call pzolopd1('N', rk, n, n, alpha, beta, a, 1, 1, desca, q, 1, 1, descq, info, 'N')
  ok = ok .and. info == -1409
  descq(9) = n

  ! PQDWHFACS: workspace query, then the run with H in WORK2
  allocate (work(2))
  call pqdwhfacs('N', n, n, 0d0, a, 1, 1, desca, q, 1, 1, descq, work, -1, descw, tau, work, -1, info, 'T')
  lwork = int(work(1))
  lwork2 = int(work(2))
  deallocate (work)
  allocate (work(lwork), work2(lwork2))
  call pqdwhfacs('N', n, n, 0d0, a, 1, 1, desca, q, 1, 1, descq, work, lwork, descw, tau, work2, lwork2 - 1, &
                 info, 'T')
  ok = ok .and. info == -16
  call pqdwhfacs('N', n, n, 0d0, a, 1, 1, desca, q, 1, 1, descq, work, lwork, descw, tau, work2, lwork2, info, 'T')
  h = reshape(work2(1:n * n), (/n, n/))
  rel = norm2(a - matmul(q, h)) / norm2(a)
  orth = norm2(matmul(transpose(q), q) - eye(n)) / sqrt(dble(n))
  write (*, '(a,i6,a,16a1,a,es10.3,a,es10.3)') 'pqdwhfacs info=', info, ' schedule=', zolo_last_info%schedule, &
    ' bwd=', rel, ' orth=', orth
  ok = ok .and. info == 0 .and. rel < 1d-13 .and. orth < 1d-13

  if (ok) then
    write (*, '(a)') 'PASS'
  else
    write (*, '(a)') 'FAIL'
  end if

contains

  function eye(m) result(e)
    integer, intent(in) :: m
    double precision :: e(m, m)
    integer :: k
    e = 0d0
    do k = 1, m
      e(k, k) = 1d0
    end do
  end function eye

  subroutine checks(u, relerr, orthe)
    double precision, intent(in) :: u(n, n)
    double precision, intent(out) :: relerr, orthe
    double precision :: t(n, n)
    t = matmul(transpose(u), a)
    t = 0.5d0 * (t + transpose(t))
    relerr = norm2(a - matmul(u, t)) / norm2(a)
    orthe = norm2(matmul(transpose(u), u) - eye(n)) / sqrt(dble(n))
  end subroutine checks

  ! A = Q1 diag(s) Q2^T with Q1, Q2 products of n Householder reflectors from an LCG.
  subroutine make_matrix(m, sv, out)
    integer, intent(in) :: m
    double precision, intent(in) :: sv(m)
    double precision, intent(out) :: out(m, m)
    double precision :: q1(m, m), q2(m, m)
    integer(8) :: state
    state = 20260101_8
    call haar_like(m, q1, state)
    call haar_like(m, q2, state)
    out = matmul(q1 * spread(sv, 1, m), transpose(q2))
  end subroutine make_matrix

  subroutine haar_like(m, qq, state)
    integer, intent(in) :: m
    double precision, intent(out) :: qq(m, m)
    integer(8), intent(inout) :: state
    double precision :: v(m), w(m), nv
    integer :: k, j
    qq = eye(m)
    do k = 1, m
      do j = 1, m
        state = mod(16807_8 * state, 2147483647_8)   ! Park-Miller
        v(j) = dble(state) / 2147483647d0 - 0.5d0
      end do
      nv = norm2(v)
      v = v / nv
      w = matmul(qq, v)
      do j = 1, m
        qq(:, j) = qq(:, j) - 2d0 * w * v(j)
      end do
    end do
  end subroutine haar_like
end program test_pzolopd1
