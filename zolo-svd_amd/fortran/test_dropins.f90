!> Every Fortran drop-in not covered by test_pzolopd1.f90, called once through
!> its reference argument list and checked on the host:
!>   ZOLOCHOL (zolochol.f90:2) and ZOLOQR (zoloqr.f90:2): Z (X^T X + c I) = -a X,
!>     and the two agree (Q1 Q2^T / sqrt(c) = X (X^T X + c I)^{-1});
!>   PDNORMEST (pdnormest.f90:2) on a submatrix (IA = 3, JA = 2) of a larger
!>     array: the 2-norm of A = Q1 diag(s) Q2^T is s(1);
!>   MPDGEQRF / MPDORGQR (mpdgeqrf.f:5, mpdorgqr.f:5) on the stack [X; sqrt(c) I]
!>     at IA = 4, JA = 3: Q^T Q = I and Q R = the stack;
!>   FormX2 (FormX2.f90:2): X+ = (X + Z)/FONE and ERR = ||X+ - X||_F;
!>   PZOLOPD2 (pzolopd2.f90:2) and PZOLOPD1 with A and Q at IA = JA = IQ = JQ = 2
!>     of larger arrays: the same polar factor as the contiguous PZOLOPD1 call.
!> Prints one line per check and PASS or FAIL.
program test_dropins
  use zolo_capi
  use auxil_zolo
  implicit none
  integer, parameter :: n = 96, big = n + 7
  double precision :: x(n, n), z1(n, n), z2(n, n), z3(n, n), w(n), coeff(24), s(n), con, c, rel
  double precision :: a(n, n), abig(big, big), qbig(big, big), q0(n, n), stack(2 * n + 9, n + 4), r(n, n)
  double precision :: tau(n), work(4 * n), cond, fone, err, alpha, beta, xs(n, n)
  integer :: rk, info, cnt, i, j, desc(9), descb(9), descs(9)
  logical :: ok
  external :: zolochol, zoloqr, pdnormest, mpdgeqrf, mpdorgqr, formx2, pzolopd1, pzolopd2

  ok = .true.
  do i = 1, n
    s(i) = 10d0 ** (-4d0 * dble(i - 1) / dble(n - 1))
  end do
  call make_matrix(n, s, x)

  ! ---- ZOLOCHOL and ZOLOQR, shift ip = 2 of r = 3 at CON = 1e4
  con = 1d4
  rk = 3
  call compute_coeff(con, epsilon(1d0), rk, coeff)
  c = coeff(3)
  xs = x * 1d4   ! iterate X = A / (alpha beta) of singular values in [1, 1e4]
  call zolochol('N', n, 64, 1d0, xs, z1, 0, work, coeff, rk, 2, info)
  z3 = matmul(z1, matmul(transpose(xs), xs) + c * eye(n)) + coeff(2 * rk + 2) * xs
  rel = norm2(z3) / norm2(coeff(2 * rk + 2) * xs)
  write (*, '(a,i4,a,es10.3)') 'zolochol info=', info, ' residual=', rel
  ok = ok .and. info == 0 .and. rel < 1d-10
  call zoloqr(n, n, 64, xs, z2, 0, work, tau, coeff, rk, 2, info, 0)
  rel = norm2(z2 - z1) / norm2(z1)
  write (*, '(a,i4,a,es10.3)') 'zoloqr info=', info, ' |Zqr - Zchol|/|Zchol|=', rel
  ok = ok .and. info == 0 .and. rel < 1d-10
  ! SYM = 'S' on a symmetric X: the same as 'N'
  a = 0.5d0 * (xs + transpose(xs))
  call zolochol('N', n, 64, 1d0, a, z1, 0, work, coeff, rk, 1, info)
  call zolochol('S', n, 64, 1d0, a, z2, 0, work, coeff, rk, 1, info)
  rel = norm2(z2 - z1) / norm2(z1)
  write (*, '(a,i4,a,es10.3)') 'zolochol sym=S info=', info, ' |ZS - ZN|/|ZN|=', rel
  ok = ok .and. info == 0 .and. rel < 1d-10

  ! ---- PDNORMEST on a submatrix
  abig = 7d0
  abig(3:n + 2, 2:n + 1) = x
  descb = (/1, 0, big, big, 64, 64, 0, 0, big/)
  call pdnormest(n, n, abig, 3, 2, descb, 1d-10, cond, cnt)
  rel = abs(cond - s(1)) / s(1)
  write (*, '(a,i4,a,es12.5,a,es10.3)') 'pdnormest cnt=', cnt, ' cond=', cond, ' |cond - s1|/s1=', rel
  ok = ok .and. cnt > 0 .and. rel < 1d-6

  ! ---- MPDGEQRF / MPDORGQR on [X; sqrt(c) I] at IA = 4, JA = 3
  stack = 3d0
  stack(4:n + 3, 3:n + 2) = x
  stack(n + 4:2 * n + 3, 3:n + 2) = sqrt(c) * eye(n)
  descs = (/1, 0, 2 * n + 9, n + 4, 64, 64, 0, 0, 2 * n + 9/)
  call mpdgeqrf(2 * n, n, stack, 4, 3, descs, tau, work, size(work), info)
  ok = ok .and. info == 0
  r = 0d0
  do j = 1, n
    r(1:j, j) = stack(4:j + 3, j + 2)
  end do
  call mpdorgqr(2 * n, n, n, stack, 4, 3, descs, tau, work, size(work), info)
  ok = ok .and. info == 0
  z1 = matmul(transpose(stack(4:2 * n + 3, 3:n + 2)), stack(4:2 * n + 3, 3:n + 2)) - eye(n)
  rel = norm2(matmul(stack(4:n + 3, 3:n + 2), r) - x) / norm2(x)
  write (*, '(a,i4,a,es10.3,a,es10.3)') 'mpdgeqrf/mpdorgqr info=', info, ' |Q^TQ - I|=', norm2(z1), &
    ' |Q1 R - X|/|X|=', rel
  rel = max(rel, norm2(matmul(stack(n + 4:2 * n + 3, 3:n + 2), r) - sqrt(c) * eye(n)) / (sqrt(c) * sqrt(dble(n))))
  ok = ok .and. norm2(z1) < 1d-13 * n .and. rel < 1d-13 * n
  ! the columns outside the submatrix are untouched
  ok = ok .and. all(stack(:, 1:2) == 3d0) .and. all(stack(1:3, :) == 3d0)

  ! ---- FormX2
  z1 = x
  z2 = 1d-3 * transpose(x)
  z3 = x
  fone = 1.25d0
  call formx2(n, 64, fone, err, z1, z2, 0, work, info, 0, 1)
  rel = maxval(abs(z1 - (z3 + 1d-3 * transpose(x)) * (1d0 / fone)))
  write (*, '(a,i4,a,es10.3,a,es10.3)') 'formx2 info=', info, ' max|X+ - (X+Z)/FONE|=', rel, ' |ERR - ref|=', &
    abs(err - norm2(z1 - z3))
  ok = ok .and. info == 0 .and. rel == 0d0 .and. abs(err - norm2(z1 - z3)) <= 1d-14 * norm2(z1 - z3)
  ok = ok .and. all(z2 == z1)

  ! ---- PZOLOPD1 / PZOLOPD2 with A and Q inside larger arrays
  call zolo_bounds(0, n, x, n, alpha, beta, info)
  rk = choosem(1d0 / beta)
  desc = (/1, 0, n, n, 64, 64, 0, 0, n/)
  call pzolopd1('N', rk, n, n, alpha, beta, x, 1, 1, desc, q0, 1, 1, desc, info, 'N')
  ok = ok .and. info == 0
  abig = 7d0
  abig(2:n + 1, 2:n + 1) = x
  qbig = 5d0
  call pzolopd2('N', rk, n, n, alpha, beta, abig, 2, 2, descb, qbig, 2, 2, descb, info, 'N')
  rel = norm2(qbig(2:n + 1, 2:n + 1) - q0)
  write (*, '(a,i4,a,i3,a,es10.3)') 'pzolopd2 (IA=JA=2) info=', info, ' restarts=', zolo_last_info%restarts, &
    ' |Q - Q(pzolopd1)|=', rel
  ok = ok .and. info == 0 .and. rel < 1d-13 .and. all(qbig(1, :) == 5d0) .and. all(qbig(:, 1) == 5d0)
  qbig = 5d0
  call pzolopd1('N', rk, n, n, alpha, beta, abig, 2, 2, descb, qbig, 2, 2, descb, info, 'N')
  rel = norm2(qbig(2:n + 1, 2:n + 1) - q0)
  write (*, '(a,i4,a,es10.3)') 'pzolopd1 (IA=JA=2) info=', info, ' |Q - Q(contiguous)|=', rel
  ok = ok .and. info == 0 .and. rel == 0d0 .and. all(qbig(big, :) == 5d0)

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

  ! A = Q1 diag(s) Q2^T with Q1, Q2 products of Householder reflectors from an LCG.
  subroutine make_matrix(m, sv, out)
    integer, intent(in) :: m
    double precision, intent(in) :: sv(m)
    double precision, intent(out) :: out(m, m)
    double precision :: q1(m, m), q2(m, m)
    integer(8) :: state
    state = 20260102_8
    call refl_product(m, q1, state)
    call refl_product(m, q2, state)
    out = matmul(q1 * spread(sv, 1, m), transpose(q2))
  end subroutine make_matrix

  subroutine refl_product(m, qq, state)
    integer, intent(in) :: m
    double precision, intent(out) :: qq(m, m)
    integer(8), intent(inout) :: state
    double precision :: v(m), wv(m)
    integer :: k, jj
    qq = eye(m)
    do k = 1, m
      do jj = 1, m
        state = mod(16807_8 * state, 2147483647_8)   ! Park-Miller
        v(jj) = dble(state) / 2147483647d0 - 0.5d0
      end do
      v = v / norm2(v)
      wv = matmul(qq, v)
      do jj = 1, m
        qq(:, jj) = qq(:, jj) - 2d0 * wv * v(jj)
      end do
    end do
  end subroutine refl_product
end program test_dropins
