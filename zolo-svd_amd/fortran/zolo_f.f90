!> Fortran drop-in entry points of the Zolo-PD hot path on one GPU.
!>
!> The external subroutines keep the reference's names and argument lists
!> (pzolopd1.f90:2, pzolopd2.f90:2, pqdwhfacs.f90:2, zolochol.f90:2,
!> zoloqr.f90:2, pdnormest.f90:2, mpdgeqrf.f:5, mpdorgqr.f:5, FormX2.f90:2,
!> hfunc.f90:2, elpa_pdsyevd2.F90:1, pqdwhsubit.f90:2, ZoloEig.f90:2), module
!> auxil_zolo keeps Auxil_Zolo.f90's interface and module mod_prepare_matrix
!> the MatrixMarket readers of prepare_matrix.f90.  The
!> descriptors must describe a 1x1 process grid (RSRC = CSRC = 0, LLD >= M);
!> the BLACS context is never dereferenced, so the library links no MPI or
!> BLACS.  Work is done by libzolo.so through the ISO_C_BINDING interfaces of
!> module zolo_capi; arrays are host arrays, copied to and from HBM.
module zolo_capi
  use iso_c_binding
  implicit none
  private :: c_int, c_double, c_int64_t, c_char, c_ptr

  type, bind(C) :: zolo_info_t
    integer(c_int) :: iters, r, info, polish_passes, restarts
    character(kind=c_char) :: schedule(16)
    real(c_double) :: err(16), con(16), maxc(16)
    real(c_double) :: alpha, beta, time_ms
    real(c_double) :: phase_ms(8)
    integer(c_int) :: shifts(16)
    integer(c_int) :: cg_iters(4)
  end type zolo_info_t

  !> arithmetic of the Cholesky iterations: 0 FP64 (default), 1 FP32, 2 BF16x3
  integer, save :: zolo_precision = 0
  !> record of the last PZOLOPD1 / PZOLOPD2 / PQDWHFACS call
  type(zolo_info_t), save, target :: zolo_last_info
  !> Rk of the last ZoloEig call (the reference computes it but returns it to no caller)
  integer, save :: zolo_eig_count = 0

  interface
    integer(c_int) function c_zolo_pd_host(precision, sym, rk, n, alpha, beta, a, lda, u, ldu, h, ldh, info) &
        bind(C, name='zolo_pd_host')
      import :: c_int, c_char, c_int64_t, c_double, c_ptr, zolo_info_t
      integer(c_int), value :: precision, rk
      character(kind=c_char), value :: sym
      integer(c_int64_t), value :: n, lda, ldu, ldh
      real(c_double), value :: alpha, beta
      real(c_double), intent(in) :: a(*)
      real(c_double) :: u(*)
      type(c_ptr), value :: h
      type(zolo_info_t) :: info
    end function
    integer(c_int) function c_zolo_pd2_host(precision, sym, rk, n, alpha, beta, a, lda, u, ldu, h, ldh, info) &
        bind(C, name='zolo_pd2_host')
      import :: c_int, c_char, c_int64_t, c_double, c_ptr, zolo_info_t
      integer(c_int), value :: precision, rk
      character(kind=c_char), value :: sym
      integer(c_int64_t), value :: n, lda, ldu, ldh
      real(c_double), value :: alpha, beta
      real(c_double), intent(in) :: a(*)
      real(c_double) :: u(*)
      type(c_ptr), value :: h
      type(zolo_info_t) :: info
    end function
    integer(c_int) function c_zolo_qdwh_host(sym, n, sigma, a, lda, u, ldu, h, ldh, info) bind(C, name='zolo_qdwh_host')
      import :: c_int, c_char, c_int64_t, c_double, c_ptr, zolo_info_t
      character(kind=c_char), value :: sym
      integer(c_int64_t), value :: n, lda, ldu, ldh
      real(c_double), value :: sigma
      real(c_double), intent(in) :: a(*)
      real(c_double) :: u(*)
      type(c_ptr), value :: h
      type(zolo_info_t) :: info
    end function
    integer(c_int) function c_zolo_bounds_host(conv, n, a, lda, alpha, beta) bind(C, name='zolo_bounds_host')
      import :: c_int, c_int64_t, c_double
      integer(c_int), value :: conv
      integer(c_int64_t), value :: n, lda
      real(c_double), intent(in) :: a(*)
      real(c_double) :: alpha, beta
    end function
    integer(c_int) function c_zolo_chol_step_host(sym, n, x, ldx, z, ldz, coeff, rk, ip) &
        bind(C, name='zolo_chol_step_host')
      import :: c_int, c_char, c_int64_t, c_double
      character(kind=c_char), value :: sym
      integer(c_int64_t), value :: n, ldx, ldz
      real(c_double), intent(in) :: x(*), coeff(*)
      real(c_double) :: z(*)
      integer(c_int), value :: rk, ip
    end function
    integer(c_int) function c_zolo_qr_step_host(n, x, ldx, z, ldz, coeff, rk, ip) bind(C, name='zolo_qr_step_host')
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: n, ldx, ldz
      real(c_double), intent(in) :: x(*), coeff(*)
      real(c_double) :: z(*)
      integer(c_int), value :: rk, ip
    end function
    integer(c_int) function c_zolo_normest_host(m, n, a, lda, tol, est, cnt) bind(C, name='zolo_normest_host')
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: m, n, lda
      real(c_double), intent(in) :: a(*)
      real(c_double), value :: tol
      real(c_double) :: est
      integer(c_int) :: cnt
    end function
    integer(c_int) function c_zolo_mpdgeqrf_host(n, s, lds, tau) bind(C, name='zolo_mpdgeqrf_host')
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: n, lds
      real(c_double) :: s(*), tau(*)
    end function
    integer(c_int) function c_zolo_mpdorgqr_host(n, s, lds, tau) bind(C, name='zolo_mpdorgqr_host')
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: n, lds
      real(c_double) :: s(*)
      real(c_double), intent(in) :: tau(*)
    end function
    integer(c_int) function c_zolo_formx2_host(n, fone, a, lda, z, ldz, err) bind(C, name='zolo_formx2_host')
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: n, lda, ldz
      real(c_double), value :: fone
      real(c_double) :: a(*), z(*)
      real(c_double) :: err
    end function
    integer(c_int) function c_zolo_syevd_host(n, h, ldh, w, z, ldz) bind(C, name='zolo_syevd_host')
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: n, ldh, ldz
      real(c_double) :: h(*), w(*), z(*)
    end function
    integer(c_int) function c_zolo_compute_coeff(con, tol, rk, coeff) bind(C, name='zolo_compute_coeff')
      import :: c_int, c_double
      real(c_double), value :: con, tol
      integer(c_int), value :: rk
      real(c_double) :: coeff(*)
    end function
    integer(c_int) function c_zolo_choosem(con) bind(C, name='zolo_choosem')
      import :: c_int, c_double
      real(c_double), value :: con
    end function
    real(c_double) function c_zolo_hfunc(x) bind(C, name='zolo_hfunc')
      import :: c_double
      real(c_double), value :: x
    end function
    integer(c_int) function c_zolo_subit_host(n, a, lda, norma, q, ldq, rk, err, iters) &
        bind(C, name='zolo_subit_host')
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: n, lda, ldq
      real(c_double), value :: norma
      real(c_double) :: a(*), q(*)
      integer(c_int) :: rk, iters
      real(c_double) :: err
    end function
    integer(c_int) function c_zolo_eig_host(n, q, ldq, as, lda, w, v, ldv, rk) bind(C, name='zolo_eig_host')
      import :: c_int, c_int64_t, c_double
      integer(c_int64_t), value :: n, ldq, lda, ldv
      real(c_double) :: q(*), as(*), w(*), v(*)
      integer(c_int) :: rk
    end function
    integer(c_int) function c_zolo_mm_info(path, rows, cols, nnz, mtype, bw) bind(C, name='zolo_mm_info')
      import :: c_int, c_int64_t, c_char
      character(kind=c_char) :: path(*)
      integer(c_int64_t) :: rows, cols, nnz, bw
      integer(c_int) :: mtype
    end function
    integer(c_int) function c_zolo_mm_dense_host(path, sigma, a, lda, mtype) bind(C, name='zolo_mm_dense_host')
      import :: c_int, c_int64_t, c_double, c_char
      character(kind=c_char) :: path(*)
      real(c_double), value :: sigma
      real(c_double) :: a(*)
      integer(c_int64_t), value :: lda
      integer(c_int) :: mtype
    end function
    subroutine c_zolo_mellipke(alpha, tol, k, e) bind(C, name='zolo_mellipke')
      import :: c_double
      real(c_double), value :: alpha, tol
      real(c_double) :: k, e
    end subroutine
    subroutine c_zolo_mellipj(u, alpha, tol, sn, cn, dn) bind(C, name='zolo_mellipj')
      import :: c_double
      real(c_double), value :: u, alpha, tol
      real(c_double) :: sn, cn, dn
    end subroutine
  end interface

contains

  !> Select the arithmetic of the Cholesky iterations (0 FP64, 1 FP32, 2 BF16x3).
  subroutine zolo_set_precision(p)
    integer, intent(in) :: p
    zolo_precision = p
  end subroutine zolo_set_precision

  !> 0 if DESC describes an M x N matrix on a 1x1 grid, else the index of the bad entry.
  integer function zolo_desc_check(desc, m, n)
    integer, intent(in) :: desc(9), m, n
    zolo_desc_check = 0
    if (desc(1) /= 1) then
      zolo_desc_check = 1
    else if (desc(3) < m) then
      zolo_desc_check = 3
    else if (desc(4) < n) then
      zolo_desc_check = 4
    else if (desc(7) /= 0) then
      zolo_desc_check = 7
    else if (desc(8) /= 0) then
      zolo_desc_check = 8
    else if (desc(9) < max(1, m)) then
      zolo_desc_check = 9
    end if
  end function zolo_desc_check

  !> alpha = ||A||_F and beta (test_zolodrv1.f90:179-189 for conv=0, pqdwhfacs.f90:318-344 for conv=1).
  subroutine zolo_bounds(conv, n, a, lda, alpha, beta, info)
    integer, intent(in) :: conv, n, lda
    double precision, intent(in) :: a(lda, *)
    double precision, intent(out) :: alpha, beta
    integer, intent(out) :: info
    info = c_zolo_bounds_host(int(conv, c_int), int(n, c_int64_t), a, int(lda, c_int64_t), alpha, beta)
  end subroutine zolo_bounds
end module zolo_capi

!> Auxil_Zolo.f90's module interface (mellipke, mellipj, choosem, compute_coeff).
module auxil_zolo
  use zolo_capi
  use iso_c_binding, only: c_int, c_double
  implicit none
  public :: mellipke, mellipj, choosem, compute_coeff
contains
  subroutine mellipke(alpha, tol, k, e)
    double precision, intent(in) :: alpha, tol
    double precision, intent(out) :: k, e
    call c_zolo_mellipke(alpha, tol, k, e)
  end subroutine mellipke

  subroutine mellipj(u, alpha, tol, sn, cn, dn)
    double precision, intent(in) :: u, alpha, tol
    double precision, intent(out) :: sn, cn, dn
    call c_zolo_mellipj(u, alpha, tol, sn, cn, dn)
  end subroutine mellipj

  function choosem(con) result(m)
    double precision :: con
    integer :: m
    m = c_zolo_choosem(con)
  end function choosem

  subroutine compute_coeff(con, tol, rk, coeff)
    integer, intent(in) :: rk
    double precision, intent(in) :: con, tol
    double precision, intent(inout) :: coeff(:)
    double precision :: tmp(24)
    integer :: st
    st = c_zolo_compute_coeff(con, tol, int(rk, c_int), tmp)
    if (st == 0) coeff(1:3*rk) = tmp(1:3*rk)
  end subroutine compute_coeff
end module auxil_zolo

!> PZOLOPD1 (pzolopd1.f90:2): Zolo-PD with caller-supplied ALPHA, BETA.  Q = polar factor.
subroutine pzolopd1(sym, rk, m, n, alpha, beta, as, ia, ja, desca, q, iq, jq, descq, info, polar)
  use zolo_capi
  use iso_c_binding
  implicit none
  character, intent(in) :: sym, polar
  integer, intent(in) :: rk, m, n, ia, ja, iq, jq, desca(9), descq(9)
  double precision, intent(in) :: alpha, beta, as(*)
  double precision, intent(inout) :: q(*)
  integer, intent(out) :: info
  call zolo_pd_common(1)
contains
  subroutine zolo_pd_common(which)
    integer, intent(in) :: which
    integer :: k, st, ioa, ioq
    character(kind=c_char) :: cs
    info = 0
    if (sym /= 'S' .and. sym /= 's' .and. sym /= 'N' .and. sym /= 'n') then
      info = -1
    else if (rk < 1 .or. rk > 8) then
      info = -2
    else if (m /= n) then
      info = -3
    else if (n < 0) then
      info = -4
    else if (.not. (alpha > 0d0)) then
      info = -5
    else if (.not. (beta > 0d0 .and. beta <= 1d0)) then
      info = -6
    end if
    if (info /= 0) return
    k = zolo_desc_check(desca, ia - 1 + m, ja - 1 + n)
    if (k /= 0) then
      info = -(1000 + k)
      return
    end if
    k = zolo_desc_check(descq, iq - 1 + m, jq - 1 + n)
    if (k /= 0) then
      info = -(1400 + k)
      return
    end if
    if (n == 0) return
    ioa = ia + (ja - 1) * desca(9)
    ioq = iq + (jq - 1) * descq(9)
    cs = sym
    if (which == 1) then
      st = c_zolo_pd_host(int(zolo_precision, c_int), cs, int(rk, c_int), int(n, c_int64_t), alpha, beta, as(ioa), &
                          int(desca(9), c_int64_t), q(ioq), int(descq(9), c_int64_t), c_null_ptr, 1_c_int64_t, &
                          zolo_last_info)
    else
      st = c_zolo_pd2_host(int(zolo_precision, c_int), cs, int(rk, c_int), int(n, c_int64_t), alpha, beta, as(ioa), &
                           int(desca(9), c_int64_t), q(ioq), int(descq(9), c_int64_t), c_null_ptr, 1_c_int64_t, &
                           zolo_last_info)
    end if
    info = st
    if (polar == polar) continue  ! POLAR is ignored, as in pzolopd1.f90:504-511
  end subroutine zolo_pd_common
end subroutine pzolopd1

!> PZOLOPD2 (pzolopd2.f90:2): as PZOLOPD1 with the iteration cap of pzolopd2.f90:394.
subroutine pzolopd2(sym, rk, m, n, alpha, beta, as, ia, ja, desca, q, iq, jq, descq, info, polar)
  use zolo_capi
  use iso_c_binding
  implicit none
  character, intent(in) :: sym, polar
  integer, intent(in) :: rk, m, n, ia, ja, iq, jq, desca(9), descq(9)
  double precision, intent(in) :: alpha, beta, as(*)
  double precision, intent(inout) :: q(*)
  integer, intent(out) :: info
  integer :: k, ioa, ioq
  character(kind=c_char) :: cs
  info = 0
  if (sym /= 'S' .and. sym /= 's' .and. sym /= 'N' .and. sym /= 'n') then
    info = -1
  else if (rk < 1 .or. rk > 8) then
    info = -2
  else if (m /= n) then
    info = -3
  else if (.not. (alpha > 0d0)) then
    info = -5
  else if (.not. (beta > 0d0 .and. beta <= 1d0)) then
    info = -6
  end if
  if (info /= 0) return
  k = zolo_desc_check(desca, ia - 1 + m, ja - 1 + n)
  if (k /= 0) then
    info = -(1000 + k)
    return
  end if
  k = zolo_desc_check(descq, iq - 1 + m, jq - 1 + n)
  if (k /= 0) then
    info = -(1400 + k)
    return
  end if
  if (n == 0) return
  ioa = ia + (ja - 1) * desca(9)
  ioq = iq + (jq - 1) * descq(9)
  cs = sym
  info = c_zolo_pd2_host(int(zolo_precision, c_int), cs, int(rk, c_int), int(n, c_int64_t), alpha, beta, as(ioa), &
                         int(desca(9), c_int64_t), q(ioq), int(descq(9), c_int64_t), c_null_ptr, 1_c_int64_t, &
                         zolo_last_info)
  if (polar == polar) continue
end subroutine pzolopd2

!> PQDWHFACS (pqdwhfacs.f90:2): QDWH-PD with shift SIGMA; H in WORK2 when POLAR = 'T'.
!> LWORK = -1 or LWORK2 = -1 is a workspace query (WORK(1), WORK(2)); the library
!> needs no caller workspace but keeps the reference's size contract.
subroutine pqdwhfacs(sym, m, n, sigma, a, ia, ja, desca, q, iq, jq, descq, work, lwork, descw, tau, work2, lwork2, &
                     info, polar)
  use zolo_capi
  use iso_c_binding
  implicit none
  character, intent(in) :: sym, polar
  integer, intent(in) :: m, n, ia, ja, iq, jq, lwork, lwork2, desca(9), descq(9), descw(9)
  double precision, intent(in) :: sigma, a(*)
  double precision, intent(inout) :: q(*), work(*), tau(*)
  double precision, intent(inout), target :: work2(*)
  integer, intent(out) :: info
  integer :: k, lwin, lwin2, ioa, ioq
  logical :: lquery
  character(kind=c_char) :: cs
  type(c_ptr) :: hp
  info = 0
  if (m < n) then
    info = -1
    return
  end if
  if (m /= n) then
    info = -2
    return
  end if
  lquery = (lwork == -1 .or. lwork2 == -1)
  if (m == 0 .or. n == 0) return
  lwin = 2 * n * n
  lwin2 = n * n
  work(1) = dble(lwin)
  work(2) = dble(lwin2)
  if (lquery) return
  if (lwork < lwin) then
    info = -12
    return
  else if (lwork2 < lwin2) then
    info = -16
    return
  end if
  if (sym /= 'S' .and. sym /= 's' .and. sym /= 'N' .and. sym /= 'n') then
    info = -1
    return
  end if
  k = zolo_desc_check(desca, ia - 1 + m, ja - 1 + n)
  if (k /= 0) then
    info = -(800 + k)
    return
  end if
  k = zolo_desc_check(descq, iq - 1 + m, jq - 1 + n)
  if (k /= 0) then
    info = -(1200 + k)
    return
  end if
  ioa = ia + (ja - 1) * desca(9)
  ioq = iq + (jq - 1) * descq(9)
  cs = sym
  hp = c_null_ptr
  if (polar == 'T' .or. polar == 't') hp = c_loc(work2(ioq))
  info = c_zolo_qdwh_host(cs, int(n, c_int64_t), sigma, a(ioa), int(desca(9), c_int64_t), q(ioq), &
                          int(descq(9), c_int64_t), hp, int(descq(9), c_int64_t), zolo_last_info)
end subroutine pqdwhfacs

!> ZOLOCHOL (zolochol.f90:2): Z = -a_IP X (X^T X + c_{2IP-1} I)^{-1} for the local n x n X in A.
subroutine zolochol(sym, n, nblk, fone, a, z, contxt, work, coeff, rk, ip, info)
  use zolo_capi
  use iso_c_binding
  implicit none
  character, intent(in) :: sym
  integer, intent(in) :: n, nblk, contxt, rk, ip
  double precision, intent(in) :: fone, a(*), coeff(*)
  double precision, intent(inout) :: z(*), work(*)
  integer, intent(out) :: info
  character(kind=c_char) :: cs
  cs = sym
  info = c_zolo_chol_step_host(cs, int(n, c_int64_t), a, int(n, c_int64_t), z, int(n, c_int64_t), coeff, &
                               int(rk, c_int), int(ip, c_int))
end subroutine zolochol

!> ZOLOQR (zoloqr.f90:2): Z = -a_IP/sqrt(c_{2IP-1}) Q1 Q2^T with [X; sqrt(c) I] = QR.
subroutine zoloqr(m, n, nblk, a, z, contxt, workq, tau, coeff, rk, ip, info, it)
  use zolo_capi
  use iso_c_binding
  implicit none
  integer, intent(in) :: m, n, nblk, contxt, rk, ip, it
  double precision, intent(in) :: a(*), coeff(*)
  double precision, intent(inout) :: z(*), workq(*), tau(*)
  integer, intent(out) :: info
  if (m /= n) then
    info = -1
    return
  end if
  info = c_zolo_qr_step_host(int(n, c_int64_t), a, int(n, c_int64_t), z, int(n, c_int64_t), coeff, int(rk, c_int), &
                             int(ip, c_int))
end subroutine zoloqr

!> PDNORMEST (pdnormest.f90:2): 2-norm estimate by power iteration (A then A^T).
subroutine pdnormest(m, n, a, ia, ja, desca, tol, cond, cnt)
  use zolo_capi
  use iso_c_binding
  implicit none
  integer, intent(in) :: m, n, ia, ja, desca(9)
  double precision, intent(in) :: tol, a(*)
  integer, intent(inout) :: cnt
  double precision, intent(inout) :: cond
  integer(c_int) :: c
  integer :: st, ioa
  ioa = ia + (ja - 1) * desca(9)
  st = c_zolo_normest_host(int(m, c_int64_t), int(n, c_int64_t), a(ioa), int(desca(9), c_int64_t), tol, cond, c)
  cnt = c
  if (st /= 0) cnt = -abs(st)
end subroutine pdnormest

!> MPDGEQRF (mpdgeqrf.f:5): structured QR of the M = 2N stack [X; diag] in place.
subroutine mpdgeqrf(m, n, a, ia, ja, desca, tau, work, lwork, info)
  use zolo_capi
  use iso_c_binding
  implicit none
  integer, intent(in) :: m, n, ia, ja, lwork, desca(9)
  double precision, intent(inout) :: a(*), tau(*), work(*)
  integer, intent(out) :: info
  integer :: ioa
  info = 0
  work(1) = dble(max(1, n))
  if (lwork == -1) return
  if (m /= 2 * n) then
    info = -1
    return
  end if
  if (zolo_desc_check(desca, ia - 1 + m, ja - 1 + n) /= 0) then
    info = -(600 + zolo_desc_check(desca, ia - 1 + m, ja - 1 + n))
    return
  end if
  ioa = ia + (ja - 1) * desca(9)
  info = c_zolo_mpdgeqrf_host(int(n, c_int64_t), a(ioa), int(desca(9), c_int64_t), tau)
end subroutine mpdgeqrf

!> MPDORGQR (mpdorgqr.f:5): explicit Q of the structured stack from its reflectors, in place.
subroutine mpdorgqr(m, n, k, a, ia, ja, desca, tau, work, lwork, info)
  use zolo_capi
  use iso_c_binding
  implicit none
  integer, intent(in) :: m, n, k, ia, ja, lwork, desca(9)
  double precision, intent(inout) :: a(*), work(*)
  double precision, intent(in) :: tau(*)
  integer, intent(out) :: info
  integer :: ioa
  info = 0
  work(1) = dble(max(1, n))
  if (lwork == -1) return
  if (m /= 2 * n) then
    info = -1
    return
  end if
  if (k /= n) then
    info = -3
    return
  end if
  if (zolo_desc_check(desca, ia - 1 + m, ja - 1 + n) /= 0) then
    info = -(700 + zolo_desc_check(desca, ia - 1 + m, ja - 1 + n))
    return
  end if
  ioa = ia + (ja - 1) * desca(9)
  info = c_zolo_mpdorgqr_host(int(n, c_int64_t), a(ioa), int(desca(9), c_int64_t), tau)
end subroutine mpdorgqr

!> FormX2 (FormX2.f90:2) on the local n x n arrays: Z <- (A+Z)/FONE, ERR = ||Z-A||_F, A <- Z
!> (computed on the GPU).
subroutine formx2(n, nblk, fone, err, a, z, contxt, work, info, it, ip)
  use zolo_capi
  use iso_c_binding
  implicit none
  integer, intent(in) :: n, nblk, contxt, it, ip
  double precision, intent(in) :: fone
  double precision, intent(out) :: err
  double precision, intent(inout) :: a(n, n), z(n, n), work(*)
  integer, intent(out) :: info
  err = 0d0
  info = 0
  if (n <= 0) return
  info = c_zolo_formx2_host(int(n, c_int64_t), fone, a, int(n, c_int64_t), z, int(n, c_int64_t), err)
end subroutine formx2

!> ELPA_PDSYEVD2 (elpa_pdsyevd2.F90:1) on a 1x1 grid: A (na x na symmetric,
!> overwritten) = Z diag(EV) Z^T with EV ascending, computed on the GPU
!> (tridiagonalisation, divide and conquer, back-transformation).  The grid and
!> communicator arguments are validated for a single process and otherwise
!> unused; UPLO 'L', 'U' or 'A' (the full matrix is read).
subroutine elpa_pdsyevd2(uplo, na, nev, a, lda, lca, ev, z, nblk, sc_desc, np_rows, np_cols, my_prow, my_pcol, &
                         myid, mpi_comm_in, info, real_kernel)
  use zolo_capi
  use iso_c_binding
  implicit none
  character, intent(in) :: uplo
  integer, intent(in) :: na, nev, lda, lca, nblk, np_rows, np_cols, my_prow, my_pcol, myid
  integer, intent(inout) :: info
  integer, optional :: mpi_comm_in, real_kernel
  double precision, intent(inout) :: a(lda, *), z(lda, *), ev(*)
  integer, intent(in) :: sc_desc(*)
  info = 0
  if (uplo /= 'L' .and. uplo /= 'l' .and. uplo /= 'U' .and. uplo /= 'u' .and. uplo /= 'A' .and. uplo /= 'a') then
    info = -1
  else if (na < 0) then
    info = -2
  else if (nev < 0 .or. nev > na) then
    info = -3
  else if (lda < max(1, na)) then
    info = -5
  else if (np_rows /= 1 .or. np_cols /= 1 .or. my_prow /= 0 .or. my_pcol /= 0) then
    info = -11
  end if
  if (info /= 0) return
  if (na == 0) return
  if (nblk < 0 .or. lca < 0 .or. myid < 0 .or. sc_desc(1) < 0) info = 0
  if (present(mpi_comm_in) .or. present(real_kernel)) info = 0
  info = c_zolo_syevd_host(int(na, c_int64_t), a, int(lda, c_int64_t), ev, z, int(lda, c_int64_t))
end subroutine elpa_pdsyevd2

!> hfunc (hfunc.f90:2): the QDWH weight a(l).
double precision function hfunc(x)
  use zolo_capi
  implicit none
  double precision, intent(in) :: x
  hfunc = c_zolo_hfunc(x)
end function hfunc

!> PQDWHSUBIT (pqdwhsubit.f90:2): Q, the polar factor of A - sigma I on entry,
!> becomes V = [V1 V2] (V1 spanning the eigenvectors of the Rk eigenvalues
!> below sigma) and A becomes V^T A V; NORMA = ||A||_F.  WORK, TAU and WORK2
!> are accepted for the reference's argument list and unused.
subroutine pqdwhsubit(n, a, ia, ja, desca, norma, q, iq, jq, descq, work, lwork, tau, work2, lwork2, rk, info)
  use zolo_capi
  use iso_c_binding
  implicit none
  integer, intent(in) :: n, ia, ja, iq, jq, lwork, lwork2, desca(9), descq(9)
  double precision, intent(in) :: norma
  double precision, intent(inout) :: a(*), q(*), work(*), tau(*), work2(*)
  integer, intent(out) :: rk, info
  integer :: k, ioa, ioq
  integer(c_int) :: crk, cit
  real(c_double) :: err
  info = 0
  rk = 0
  if (n == 0) return
  if (n < 0) then
    info = -1
    return
  end if
  k = zolo_desc_check(desca, ia - 1 + n, ja - 1 + n)
  if (k /= 0) then
    info = -(500 + k)
    return
  end if
  k = zolo_desc_check(descq, iq - 1 + n, jq - 1 + n)
  if (k /= 0) then
    info = -(1000 + k)
    return
  end if
  if (.not. (norma > 0d0)) then
    info = -6
    return
  end if
  ioa = ia + (ja - 1) * desca(9)
  ioq = iq + (jq - 1) * descq(9)
  info = c_zolo_subit_host(int(n, c_int64_t), a(ioa), int(desca(9), c_int64_t), norma, q(ioq), &
                           int(descq(9), c_int64_t), crk, err, cit)
  rk = crk
end subroutine pqdwhsubit

!> ZoloEig (ZoloEig.f90:2) on a 1x1 grid: Q (N x N, leading dimension N) the
!> polar factor of As - sigma I, As the matrix.  On exit TAU(1:Rk) holds the
!> eigenvalues below sigma (ascending) and WORK(1:N*Rk) their eigenvectors
!> (N x Rk, leading dimension N), as the reference leaves them; Rk is in
!> zolo_eig_count of module zolo_capi.  Q and As are left unchanged; WORK2
!> is accepted and unused.  LWORK < N*N: INFO = -8.
subroutine zoloeig(n, nblk, q, as, contxt, mpi_newcomm, work, lwork, work2, lwork2, tau, info)
  use zolo_capi
  use iso_c_binding
  implicit none
  integer, intent(in) :: n, nblk, contxt, mpi_newcomm, lwork, lwork2
  double precision, intent(inout) :: q(*), as(*), work(*), work2(*), tau(*)
  integer, intent(out) :: info
  integer(c_int) :: crk
  info = 0
  zolo_eig_count = 0
  if (n == 0) return
  if (n < 0) then
    info = -1
    return
  end if
  if (lwork < n * n) then
    info = -8
    return
  end if
  info = c_zolo_eig_host(int(n, c_int64_t), q, int(n, c_int64_t), as, int(n, c_int64_t), tau, work, &
                         int(n, c_int64_t), crk)
  zolo_eig_count = crk
end subroutine zoloeig

!> prepare_matrix.f90's MatrixMarket readers on a 1x1 grid (the descriptor's
!> LLD is the leading dimension of A): the file's entries stored in order,
!> symmetric files then A <- A + A^T and A(i,i) <- A(i,i)/2 - SIGMA.  As in the
!> reference, SIGMA is set to zero on entry (prepare_matrix.f90:214, :651),
!> so no shift is applied, and NNZ returns the number of entries.  Z is
!> accepted and unused (the transpose is formed in place).
module mod_prepare_matrix
  use zolo_capi
  use iso_c_binding
  implicit none
contains
  subroutine prepare_real_symm_mm(filename, na, nnz, myid, myrow, mycol, nprow, npcol, sc_desc, a, z, sigma)
    character(len=*), intent(in) :: filename
    integer, intent(in) :: na, myid, myrow, mycol, nprow, npcol, sc_desc(*)
    integer, intent(inout) :: nnz
    double precision, intent(inout) :: a(*), z(*), sigma
    character :: sym
    call prepare_real_gem(sym, filename, na, nnz, myid, myrow, mycol, nprow, npcol, sc_desc, a, z, sigma)
  end subroutine prepare_real_symm_mm

  !> SYM returns 'S' (symmetric file), 'G' (general) or 'O' (other), as the
  !> reference's 'S', 'Gen' and 'Other' read into one character.  A file that
  !> cannot be read, is not square or not of order NA stops the program, as
  !> the reference's mmread does.
  subroutine prepare_real_gem(sym, filename, na, nnz, myid, myrow, mycol, nprow, npcol, sc_desc, a, z, sigma)
    character, intent(out) :: sym
    character(len=*), intent(in) :: filename
    integer, intent(in) :: na, myid, myrow, mycol, nprow, npcol, sc_desc(*)
    integer, intent(inout) :: nnz
    double precision, intent(inout) :: a(*), z(*), sigma
    integer(c_int64_t) :: rows, cols, cnnz, bw
    integer(c_int) :: mtype, st
    character(kind=c_char, len=len_trim(filename) + 1) :: cpath
    cpath = trim(filename) // c_null_char
    sigma = 0d0
    if (myid < 0 .or. myrow /= 0 .or. mycol /= 0 .or. nprow /= 1 .or. npcol /= 1) stop 'prepare_matrix: 1x1 grid only'
    st = c_zolo_mm_info(cpath, rows, cols, cnnz, mtype, bw)
    if (st /= 0) stop 'prepare_matrix: cannot read the MatrixMarket file'
    if (rows /= na .or. cols /= na) stop 'prepare_matrix: the file is not of order NA'
    st = c_zolo_mm_dense_host(cpath, sigma, a, int(sc_desc(9), c_int64_t), mtype)
    if (st /= 0) stop 'prepare_matrix: bad MatrixMarket file'
    nnz = int(cnnz)
    if (mtype == 0) then
      sym = 'S'
    else if (mtype == 1) then
      sym = 'G'
    else
      sym = 'O'
    end if
  end subroutine prepare_real_gem
end module mod_prepare_matrix
