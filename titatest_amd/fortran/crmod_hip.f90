!> CRMod-style forward modelling on the MI355X: reads a tomodir's elem.dat,
!> elec.dat, rho.dat and config.dat, runs the forward path through the C-ABI
!> and writes volt.dat the way the reference's wdatm does (src/wdatm.f90:55-79):
!> FP32 values in the list-directed layout of the reference's own files, or
!> 17 significant digits when CRMOD_HIP_FULL_PRECISION is set.
!>
!>   crmod_hip elem.dat elec.dat rho.dat config.dat volt.dat [swrtr [nsink [sens.dat]]]
!>   crmod_hip --format-test in.txt out.txt   (host only: format the REAL(4)
!>                                             values of in.txt, one per line)
!>
!> swrtr = 1 (default) is the 2.5D forward; 0 is 2-D (one wavenumber 0).
!> nsink > 0 adds the fictitious sink at that node (crmod.cfg lsink/nsink).
!> With a sensitivity file name (crmod.cfg's lsens and sens path), the
!> sensitivities dV/dsigma (bsens) are written one file per measurement as the
!> reference's wsens does (src/wsens.f90:40-98): the measurement number,
!> zero-padded to the digits of nanz, goes before the extension; first line
!> |sum_j s_ij| and its phase [mrad], then x z Re Im per cell, in FP32.
!> The readers follow src/relem.f90:62-151 (with the rnr repair of :209-238),
!> src/relectr.f90:46-68, src/rsigma.f90:91-189 and src/rdatm.f90:54-115.
program crmod_hip
  use iso_c_binding
  use crt_hip
  implicit none

  integer(c_int32_t) :: sanz, typanz, mb, eanz, nanz, elanz, relanz, maxn, nel
  integer(c_int32_t), allocatable, target :: typ(:), nelanz(:), selanz(:), snr(:), nrel(:,:), rnr(:), enr(:)
  integer(c_int32_t), allocatable :: strnr(:), vnr(:)
  real(c_double), allocatable, target :: sx(:), sy(:)
  complex(c_double_complex), allocatable :: sigma(:), volt(:), sigmaa(:)
  complex(c_double_complex), allocatable, target :: sens(:,:)
  real(c_double) :: kwn(30), kwnwi(30), bet, pha, pi, npi
  integer(c_int32_t) :: kwnanz, lbeta, lrandb, swrtr, nsink, code, kk, ll
  type(crt_mesh) :: mesh
  type(c_ptr) :: h
  character(len=512) :: f_elem, f_elec, f_rho, f_cfg, f_volt, f_sens, arg
  integer :: i, j, k, idum, e1, e2, ios, rc, ie1, ie2, ic, jk1, jk2, l
  real(c_double) :: rb, rp
  logical :: full

  if (command_argument_count() == 3) then
     call get_command_argument(1, arg)
     if (trim(arg) == '--format-test') then
        call get_command_argument(2, f_elem)
        call get_command_argument(3, f_volt)
        call format_test(trim(f_elem), trim(f_volt))
        stop
     end if
  end if
  if (command_argument_count() < 5) then
     write (*, '(a)') 'usage: crmod_hip elem.dat elec.dat rho.dat config.dat volt.dat [swrtr [nsink [sens.dat]]]'
     stop 1
  end if
  call get_command_argument(1, f_elem)
  call get_command_argument(2, f_elec)
  call get_command_argument(3, f_rho)
  call get_command_argument(4, f_cfg)
  call get_command_argument(5, f_volt)
  swrtr = 1
  nsink = 0
  if (command_argument_count() >= 6) then
     call get_command_argument(6, arg)
     read (arg, *) swrtr
  end if
  if (command_argument_count() >= 7) then
     call get_command_argument(7, arg)
     read (arg, *) nsink
  end if
  f_sens = ''
  if (command_argument_count() >= 8) call get_command_argument(8, f_sens)
  pi = acos(-1d0)

  ! ---- elem.dat
  open (10, file=trim(f_elem), status='old', action='read')
  read (10, *) sanz, typanz, mb
  allocate (typ(typanz), nelanz(typanz), selanz(typanz))
  read (10, *) (typ(i), nelanz(i), selanz(i), i = 1, typanz)
  maxn = maxval(selanz)
  elanz = 0
  relanz = 0
  do i = 1, typanz
     if (typ(i) > 10) then
        relanz = relanz + nelanz(i)
     else
        elanz = elanz + nelanz(i)
     end if
  end do
  nel = elanz + relanz
  allocate (snr(sanz), sx(sanz), sy(sanz), nrel(nel, maxn), rnr(max(relanz, 1)))
  read (10, *) (snr(i), sx(i), sy(i), i = 1, sanz)
  nrel = 0
  idum = 0
  do i = 1, typanz
     do j = 1, nelanz(i)
        read (10, *) (nrel(idum + j, k), k = 1, selanz(i))
     end do
     idum = idum + nelanz(i)
  end do
  if (relanz > 0) read (10, *) (rnr(i), i = 1, relanz)
  close (10)
  ! link each boundary element to the area element holding both its nodes
  do i = 1, relanz
     do k = 1, elanz
        jk1 = 0
        jk2 = 0
        do l = 1, selanz(1)
           if (nrel(k, l) == nrel(elanz + i, 1)) jk1 = 1
           if (nrel(k, l) == nrel(elanz + i, 2)) jk2 = 1
        end do
        if (jk1 == 1 .and. jk2 == 1) rnr(i) = k
     end do
  end do

  ! ---- elec.dat
  open (11, file=trim(f_elec), status='old', action='read')
  read (11, *) eanz
  allocate (enr(eanz))
  do i = 1, eanz
     read (11, *) enr(i)
  end do
  close (11)

  ! ---- rho.dat: sigma = (cos(phi)/|rho|, -sin(phi)/|rho|), phi in mrad
  open (12, file=trim(f_rho), status='old', action='read')
  read (12, *) idum
  if (idum /= elanz) then
     write (*, '(a)') 'rho.dat does not match the mesh (errnr 47)'
     stop 47
  end if
  allocate (sigma(elanz))
  do i = 1, elanz
     read (12, *) rb, rp
     rp = 1d-3 * rp
     sigma(i) = dcmplx(dcos(rp) / rb, -dsin(rp) / rb)
  end do
  close (12)

  ! ---- config.dat (A*10^4+B  M*10^4+N, or A B M N)
  open (13, file=trim(f_cfg), status='old', action='read')
  read (13, *) nanz
  allocate (strnr(nanz), vnr(nanz), volt(nanz), sigmaa(nanz))
  read (13, *) idum
  backspace (13)
  do i = 1, nanz
     if (idum >= 10000) then
        read (13, *) strnr(i), vnr(i)
     else
        read (13, *) e1, e2, ie1, ie2
        strnr(i) = e1 * 10000 + e2
        vnr(i) = ie1 * 10000 + ie2
     end if
  end do
  close (13)

  ! ---- the forward path on the device
  mesh%sanz = sanz
  mesh%typanz = typanz
  mesh%mb = mb
  mesh%eanz = eanz
  mesh%typ = c_loc(typ)
  mesh%nelanz = c_loc(nelanz)
  mesh%selanz = c_loc(selanz)
  mesh%nrel = c_loc(nrel)
  mesh%rnr = c_loc(rnr)
  mesh%snr = c_loc(snr)
  mesh%sx = c_loc(sx)
  mesh%sy = c_loc(sy)
  mesh%enr = c_loc(enr)
  mesh%lsytop = 0
  mesh%lsink = merge(1, 0, nsink > 0)
  mesh%nsink = nsink
  rc = crt_create(h, mesh)
  call check(rc, 'crt_create')
  if (swrtr == 1) then
     rc = crt_rwaven(h, nanz, strnr, vnr, kwnanz, kwn, kwnwi)
     call check(rc, 'crt_rwaven')
  else
     kwn(1) = 0d0
     kwnwi(1) = 1d0
     rc = crt_set_wavenumbers(h, 1, kwn, kwnwi, 0)
     call check(rc, 'crt_set_wavenumbers')
  end if
  rc = crt_precal(h, lbeta, lrandb)
  call check(rc, 'crt_precal')
  rc = crt_forward(h, sigma, c_null_ptr, c_null_ptr)
  call check(rc, 'crt_forward')
  rc = crt_voltages(h, nanz, strnr, vnr, 0, c_null_ptr, volt, sigmaa)
  call check(rc, 'crt_voltages')
  if (len_trim(f_sens) > 0) then
     allocate (sens(nanz, elanz))
     rc = crt_sensitivities(h, 0, c_loc(sens))
     call check(rc, 'crt_sensitivities')
     call write_sens(trim(f_sens))
  end if
  rc = crt_destroy(h)

  ! ---- volt.dat as wdatm writes it (polarity swap when the phase is near pi)
  call get_environment_variable('CRMOD_HIP_FULL_PRECISION', arg, status=ios)
  full = ios == 0
  open (14, file=trim(f_volt), status='replace', action='write')
  if (full) then
     write (14, *) nanz
  else
     write (14, '(a)') list_int(nanz)
  end if
  do i = 1, nanz
     bet = abs(sigmaa(i))
     pha = atan2(dimag(sigmaa(i)), dble(sigmaa(i)))
     npi = dnint(pha / pi) * pi
     if (dabs(npi) > 1d-12) then
        pha = pha - npi
        ie1 = mod(vnr(i), 10000)
        ie2 = (vnr(i) - ie1) / 10000
        vnr(i) = ie1 * 10000 + ie2
     end if
     if (full) then
        write (14, '(2I12,2ES26.17)') strnr(i), vnr(i), bet, 1d3 * pha
     else
        write (14, '(a)') list_int(strnr(i)) // list_int(vnr(i)) // list_real4(real(bet)) // &
             list_real4(real(1d3 * pha))
     end if
  end do
  close (14)
  ic = 0

contains

  ! An integer as the reference's list-directed WRITE puts it in its files
  ! (gfortran: a 12-character field, right-justified).
  function list_int(i) result(out)
    integer(c_int32_t), intent(in) :: i
    character(len=12) :: out
    write (out, '(I12)') i
  end function list_int

  ! A REAL(4) value as the reference's list-directed WRITE puts it in its
  ! files (src/wdatm.f90:76 built with gfortran): a blank, then G editing
  ! with 9 significant digits in 16 characters -- F form with four trailing
  ! blanks when the value, rounded to 9 digits, lies in [0.1, 1e9) (zero
  ! included), otherwise 1PE16.8E2.
  function list_real4(x) result(out)
    real(4), intent(in) :: x
    character(len=17) :: out
    character(len=20) :: es
    character(len=16) :: fb, fmt
    integer :: ex, lead
    if (x == 0.0) then
       ex = 0
    else
       write (es, '(ES20.8E3)') x
       read (es(len_trim(es) - 3:len_trim(es)), *) ex
    end if
    if (ex >= -1 .and. ex < 9) then
       write (fmt, '(a,i0,a)') '(F12.', 8 - ex, ')'
       write (fb, fmt) x
       ! F editing may drop the zero before the decimal point of |x| < 1
       lead = index(fb(1:12), '.')
       if (lead >= 2) then
          if (fb(lead - 1:lead - 1) == ' ' .or. fb(lead - 1:lead - 1) == '-') then
             fb(1:lead - 1) = fb(2:lead - 1) // '0'
          end if
       end if
       out = ' ' // fb(1:12) // '    '
    else
       write (fb, '(1PE16.8E2)') x
       out = ' ' // fb(1:16)
    end if
  end function list_real4

  subroutine format_test(fin, fout)
    character(len=*), intent(in) :: fin, fout
    real(4) :: v
    integer :: st
    open (21, file=fin, status='old', action='read')
    open (22, file=fout, status='replace', action='write')
    do
       read (21, *, iostat=st) v
       if (st /= 0) exit
       write (22, '(a)') list_real4(v)
    end do
    close (21)
    close (22)
  end subroutine format_test

  ! wsens (src/wsens.f90:40-98): one file per measurement, the centroids as
  ! relem sets them (src/relem.f90:134-143)
  subroutine write_sens(pattern)
    character(len=*), intent(in) :: pattern
    real(c_double), allocatable :: ex(:), ey(:)
    complex(c_double_complex) :: summe
    character(len=16) :: digits, fmt
    character(len=600) :: fname
    integer :: nd, ls, ld, ii, jj, it, iel, kk2
    allocate (ex(elanz), ey(elanz))
    ex = 0d0
    ey = 0d0
    iel = 0
    do it = 1, typanz
       do jj = 1, nelanz(it)
          iel = iel + 1
          if (typ(it) < 10) then
             do kk2 = 1, selanz(it)
                ex(iel) = ex(iel) + sx(snr(nrel(iel, kk2)))
                ey(iel) = ey(iel) + sy(snr(nrel(iel, kk2)))
             end do
             ex(iel) = ex(iel) / selanz(it)
             ey(iel) = ey(iel) / selanz(it)
          end if
       end do
    end do
    nd = int(log10(real(nanz))) + 1
    ls = index(pattern, '/', back=.true.)
    ld = ls + index(pattern(ls + 1:), '.') - 1
    write (fmt, '(a,i0,a,i0,a)') '(i', nd, '.', nd, ')'
    do ii = 1, nanz
       write (digits, fmt) ii
       fname = pattern(1:ld) // digits(1:nd) // pattern(ld + 1:min(ld + 4, len(pattern)))
       open (16, file=trim(fname), status='replace', action='write')
       summe = dcmplx(0d0)
       do jj = 1, elanz
          summe = summe + sens(ii, jj)
       end do
       if (abs(summe) > 0d0) then
          write (16, *) real(abs(summe)), real(1d3 * atan2(dimag(summe), dble(summe)))
       else
          write (16, *) 0., 0.
       end if
       do jj = 1, elanz
          write (16, *) real(ex(jj)), real(ey(jj)), real(dble(sens(ii, jj))), real(dimag(sens(ii, jj)))
       end do
       close (16)
    end do
  end subroutine write_sens

  subroutine check(rcode, where)
    integer, intent(in) :: rcode
    character(len=*), intent(in) :: where
    if (rcode /= 0) then
       code = 0
       kk = 0
       ll = 0
       write (*, '(a,a,i6)') trim(where), ': error ', rcode
       open (15, file='error.dat', status='replace')
       write (15, '(i6)') rcode
       close (15)
       stop 2
    end if
  end subroutine check

end program crmod_hip
