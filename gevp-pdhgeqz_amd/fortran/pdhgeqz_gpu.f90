! pdhgeqz_gpu.f90 -- Fortran drop-in for the PDHGEQZ / PDGGHRD call surface of NLAFET/GEVP-PDHGEQZ
! on a 1x1 process grid, backed by the MI355X library libgqz.so through ISO_C_BINDING.
!
! Exported symbols (Fortran ABI, lower case + trailing underscore): pdhgeqz_, pdgghrd_, pdtgord_,
! kkqzconf_, pilaenvx_.  Argument lists follow src/pdhgeqz/PDHGEQZ.f:15-31,
! src/ab2ht/PDGGHRD.f:13-14, src/reorder/PDTGORD.f:1-4, src/kkqz/kkqzconf.f:1 and
! src/pdhgeqz/pilaenvx.f:1.
module gqz_c_iface
  use iso_c_binding
  implicit none
  interface
    integer(c_int) function gqz_pdhgeqz_host(job, compq, compz, n, ilo, ihi, h, desch, t, desct, &
        alphar, alphai, beta, q, descq, z, descz, dwork, ldwork, iwork, liwork) &
        bind(C, name='gqz_pdhgeqz_host')
      import :: c_int, c_double, c_char
      character(kind=c_char), value :: job, compq, compz
      integer(c_int), value :: n, ilo, ihi, ldwork, liwork
      real(c_double) :: h(*), t(*), alphar(*), alphai(*), beta(*), q(*), z(*), dwork(*)
      integer(c_int) :: desch(9), desct(9), descq(9), descz(9), iwork(*)
    end function
    integer(c_int) function gqz_pdgghrd_host(compq, compz, n, ilo, ihi, a, desca, b, descb, &
        q, descq, z, descz, work, lwork) bind(C, name='gqz_pdgghrd_host')
      import :: c_int, c_double, c_char
      character(kind=c_char), value :: compq, compz
      integer(c_int), value :: n, ilo, ihi, lwork
      real(c_double) :: a(*), b(*), q(*), z(*), work(*)
      integer(c_int) :: desca(9), descb(9), descq(9), descz(9)
    end function
    integer(c_int) function gqz_pdtgord_host(wantq, wantz, select, para, n, s, descs, t, desct, q, descq, &
        z, descz, alphar, alphai, beta, m, dwork, ldwork, iwork, liwork) bind(C, name='gqz_pdtgord_host')
      import :: c_int, c_double
      integer(c_int), value :: wantq, wantz, n, ldwork, liwork
      integer(c_int) :: select(*), para(6), descs(9), desct(9), descq(9), descz(9), m, iwork(*)
      real(c_double) :: s(*), t(*), q(*), z(*), alphar(*), alphai(*), beta(*), dwork(*)
    end function
    integer(c_int) function gqz_piparmq(ispec, nh) bind(C, name='gqz_piparmq')
      import :: c_int
      integer(c_int), value :: ispec, nh
    end function
  end interface
end module gqz_c_iface

subroutine pdhgeqz(job, compq, compz, n, ilo, ihi, h, desch, t, desct, alphar, alphai, beta, &
                   q, descq, z, descz, dwork, ldwork, iwork, liwork, info)
  use gqz_c_iface
  implicit none
  character :: job, compq, compz
  integer :: n, ilo, ihi, ldwork, liwork, info
  integer :: desch(9), desct(9), descq(9), descz(9), iwork(*)
  double precision :: h(*), t(*), alphar(*), alphai(*), beta(*), q(*), z(*), dwork(*)
  info = gqz_pdhgeqz_host(job, compq, compz, n, ilo, ihi, h, desch, t, desct, alphar, alphai, beta, &
                          q, descq, z, descz, dwork, ldwork, iwork, liwork)
end subroutine pdhgeqz

subroutine pdgghrd(compq, compz, n, ilo, ihi, a, desca, b, descb, q, descq, z, descz, work, lwork, info)
  use gqz_c_iface
  implicit none
  character :: compq, compz
  integer :: n, ilo, ihi, lwork, info
  integer :: desca(9), descb(9), descq(9), descz(9)
  double precision :: a(*), b(*), q(*), z(*), work(*)
  info = gqz_pdgghrd_host(compq, compz, n, ilo, ihi, a, desca, b, descb, q, descq, z, descz, work, lwork)
end subroutine pdgghrd

subroutine pdtgord(wantq, wantz, select, para, n, s, descs, t, desct, q, descq, z, descz, &
                   alphar, alphai, beta, m, dwork, ldwork, iwork, liwork, info)
  use gqz_c_iface
  implicit none
  logical :: wantq, wantz
  integer :: n, m, ldwork, liwork, info
  integer :: select(*), para(6), descs(9), desct(9), descq(9), descz(9), iwork(*)
  double precision :: s(*), t(*), q(*), z(*), alphar(*), alphai(*), beta(*), dwork(*)
  integer(c_int) :: iq, iz
  iq = 0; iz = 0
  if (wantq) iq = 1
  if (wantz) iz = 1
  info = gqz_pdtgord_host(iq, iz, select, para, n, s, descs, t, desct, q, descq, z, descz, &
                          alphar, alphai, beta, m, dwork, ldwork, iwork, liwork)
end subroutine pdtgord

! KKQZCONF(EPS) configured the serial QZ's COMMON blocks (src/kkqz/kkqzconf.f:44-175); the GPU
! library keeps its configuration per call, so this is a compatible no-op.
subroutine kkqzconf(eps)
  implicit none
  double precision :: eps
end subroutine kkqzconf

! PILAENVX(ICTXT, ISPEC, NAME, OPTS, N1, N2, N3, N4): the PIPARMQ branch (ISPEC 50..55); for those
! specs the reference passes N = N1, ILO = N2, IHI = N3 and the active size is IHI - ILO + 1.
! ISPEC 80..85 are the reordering parameters PARA(1..6) of PDTGORD (pilaenvx.f:651-688, N2 = NB):
! windows min(ceil(N1/N2), 1) on a 1x1 grid, N2/2 eigenvalues per window, window N2, 0 %, min(N2, 32),
! N2/2.
integer function pilaenvx(ictxt, ispec, name, opts, n1, n2, n3, n4)
  use gqz_c_iface
  implicit none
  integer :: ictxt, ispec, n1, n2, n3, n4
  character(len=*) :: name, opts
  integer :: nh
  select case (ispec)
  case (80)
    pilaenvx = min((n1 + n2 - 1) / max(n2, 1), 1)
  case (81, 85)
    pilaenvx = n2 / 2
  case (82)
    pilaenvx = n2
  case (83)
    pilaenvx = 0
  case (84)
    pilaenvx = min(n2, 32)
  case default
    nh = n3 - n2 + 1
    if (nh < 1) nh = n1
    pilaenvx = gqz_piparmq(ispec, nh)
  end select
end function pilaenvx
