!> C-callable harness around the reference CRMod/CRTomo hot-path routines.
!>
!> TEST INFRASTRUCTURE ONLY.  This file is compiled together with the
!> reference's own, unmodified sources (kompab.f90, chol.f90, vre.f90, ...),
!> which stay where they lie in the reference tree; the recipe is
!> oracle/Makefile and the output goes to oracle/_ref/ (never committed).
!> It replaces the reference's main program (src/fem.f90) with entry points
!> that load the module state from caller arrays and run one stage each, so
!> that tests can read the reference's full-precision intermediate results.
!>
!> The forward loop below follows the order of src/fem.f90:328-397
!> (kompab -> randb -> randb2 -> scalab -> chol, or kompb when the factor is
!> reused, then vre and kpot = pot*fak), with the same OpenMP split over
!> wavenumbers.
module crt_ref_driver
  use iso_c_binding
  use alloci
  use femmod
  use sigmamod
  use electrmod
  use elemmod
  use wavenmod
  use datmod
  use modelmod
  use errmod
  use konvmod, only: lverb, lsytop
  use randbmod
  implicit none

contains

  !> Load mesh + electrodes (replaces relem/relectr, src/relem.f90:62-151,
  !> src/relectr.f90:46-68).  nrel is (elanz+relanz) x maxnodes, column major,
  !> 1-based node numbers, exactly the layout of elemmod::nrel.
  integer(c_int) function ref_load_mesh(c_sanz, c_typanz, c_mb, c_typ, c_nelanz, &
       c_selanz, c_nrel, c_rnr, c_snr, c_sx, c_sy, c_eanz, c_enr, c_lsytop, &
       c_lsink, c_nsink) bind(C, name='ref_load_mesh')
    integer(c_int), value :: c_sanz, c_typanz, c_mb, c_eanz, c_lsytop, c_lsink, c_nsink
    integer(c_int), intent(in) :: c_typ(c_typanz), c_nelanz(c_typanz), c_selanz(c_typanz)
    integer(c_int), intent(in) :: c_nrel(*), c_rnr(*), c_snr(c_sanz), c_enr(c_eanz)
    real(c_double), intent(in) :: c_sx(c_sanz), c_sy(c_sanz)
    integer :: i, j, nel

    call ref_free_mesh()
    sanz = c_sanz; typanz = c_typanz; mb = c_mb
    allocate(typ(typanz), nelanz(typanz), selanz(typanz))
    typ = c_typ; nelanz = c_nelanz; selanz = c_selanz
    max_nr_element_nodes = maxval(selanz)
    elanz = 0; relanz = 0
    do i = 1, typanz
       if (typ(i) > 10) then
          relanz = relanz + nelanz(i)
       else
          elanz = elanz + nelanz(i)
       end if
    end do
    nel = elanz + relanz
    allocate(sx(sanz), sy(sanz), snr(sanz))
    sx = c_sx; sy = c_sy; snr = c_snr
    allocate(nrel(nel, max_nr_element_nodes), rnr(max(relanz,1)))
    nrel = 0
    do j = 1, max_nr_element_nodes
       do i = 1, nel
          nrel(i, j) = c_nrel(i + (j-1)*nel)
       end do
    end do
    if (relanz > 0) rnr(1:relanz) = c_rnr(1:relanz)
    eanz = c_eanz
    allocate(enr(eanz))
    enr = c_enr
    lsytop = (c_lsytop /= 0)
    lsink = (c_lsink /= 0)
    nsink = c_nsink
    lrandb2 = .false.
    rwdanz = 0
    nr_elec_capacitances = 0
    lsr = .false.
    ldc = .false.
    lverb = .false.
    manz = elanz
    allocate(mnr(elanz))
    do i = 1, elanz
       mnr(i) = i
    end do
    ref_load_mesh = 0
  end function ref_load_mesh

  subroutine ref_free_mesh()
    if (allocated(typ)) deallocate(typ, nelanz, selanz)
    if (allocated(sx)) deallocate(sx, sy, snr)
    if (allocated(nrel)) deallocate(nrel)
    if (allocated(rnr)) deallocate(rnr)
    if (allocated(enr)) deallocate(enr)
    if (allocated(mnr)) deallocate(mnr)
    if (allocated(elbg)) deallocate(elbg)
    if (allocated(relbg)) deallocate(relbg)
    if (allocated(kg)) deallocate(kg)
    if (allocated(kpot)) deallocate(kpot)
    if (allocated(hpot)) deallocate(hpot)
    if (allocated(kpotdc)) deallocate(kpotdc)
    if (allocated(hpotdc)) deallocate(hpotdc)
  end subroutine ref_free_mesh

  !> Measurement configurations (replaces rdatm, src/rdatm.f90:54-115).
  integer(c_int) function ref_load_data(c_nanz, c_strnr, c_vnr) bind(C, name='ref_load_data')
    integer(c_int), value :: c_nanz
    integer(c_int), intent(in) :: c_strnr(c_nanz), c_vnr(c_nanz)
    if (allocated(strnr)) deallocate(strnr, strom, volt, sigmaa, kfak, vnr)
    nanz = c_nanz
    allocate(strnr(nanz), strom(nanz), volt(nanz), sigmaa(nanz), kfak(nanz), vnr(nanz))
    strnr = c_strnr; vnr = c_vnr
    strom = 1d0
    kfak = 1d0
    ref_load_data = 0
  end function ref_load_data

  !> Prescribed boundary values (replaces rrandb, src/rrandb.f90): rwdnr/rwd
  !> for randb2, rwddc = their real parts for randbdc2; c_n = 0 switches
  !> lrandb2 off.
  integer(c_int) function ref_set_randb2(c_n, c_rwdnr, c_rwd) bind(C, name='ref_set_randb2')
    integer(c_int), value :: c_n
    integer(c_int), intent(in) :: c_rwdnr(*)
    complex(c_double_complex), intent(in) :: c_rwd(*)
    if (allocated(rwdnr)) deallocate(rwdnr)
    if (allocated(rwd)) deallocate(rwd)
    if (allocated(rwddc)) deallocate(rwddc)
    rwdanz = c_n
    lrandb2 = (c_n > 0)
    if (c_n > 0) then
       allocate(rwdnr(c_n), rwd(c_n), rwddc(c_n))
       rwdnr = c_rwdnr(1:c_n)
       rwd = c_rwd(1:c_n)
       rwddc = dble(rwd)
    end if
    ref_set_randb2 = 0
  end function ref_set_randb2

  !> Run the reference rwaven (src/rwaven.f90:22-171); arrays sized >= 30.
  integer(c_int) function ref_rwaven(c_kwnanz, c_kwn, c_kwnwi, c_amin, c_amax) bind(C, name='ref_rwaven')
    integer(c_int), intent(out) :: c_kwnanz
    real(c_double), intent(out) :: c_kwn(30), c_kwnwi(30), c_amin, c_amax
    if (allocated(kwn)) deallocate(kwn, kwnwi)
    swrtr = 1
    call rwaven()
    ref_rwaven = errnr
    if (errnr /= 0) return
    c_kwnanz = kwnanz
    c_kwn(1:kwnanz) = kwn
    c_kwnwi(1:kwnanz) = kwnwi
    c_amin = amin
    c_amax = amax
  end function ref_rwaven

  !> Set wavenumbers directly (2D mode: c_swrtr=0, one wavenumber 0).
  integer(c_int) function ref_set_waven(c_kwnanz, c_kwn, c_kwnwi, c_swrtr) bind(C, name='ref_set_waven')
    integer(c_int), value :: c_kwnanz, c_swrtr
    real(c_double), intent(in) :: c_kwn(c_kwnanz), c_kwnwi(c_kwnanz)
    if (allocated(kwn)) deallocate(kwn, kwnwi)
    kwnanz = c_kwnanz
    swrtr = c_swrtr
    allocate(kwn(kwnanz), kwnwi(kwnanz))
    kwn = c_kwn; kwnwi = c_kwnwi
    ref_set_waven = 0
  end function ref_set_waven

  !> Run the reference precal (src/precal.f90:13-222) and copy elbg/relbg/kg out.
  !> elbg_out: elanz x (maxnodes*(maxnodes+1)/2) x kwnanz; relbg_out: relanz x same;
  !> kg_out: relanz x eanz x kwnanz (all column major).
  integer(c_int) function ref_precal(c_lbeta, c_lrandb, c_elbg, c_relbg, c_kg, c_sytop) bind(C, name='ref_precal')
    integer(c_int), intent(out) :: c_lbeta, c_lrandb
    real(c_double), intent(out) :: c_elbg(*), c_relbg(*), c_kg(*), c_sytop
    integer :: nimn, ntot
    if (allocated(elbg)) deallocate(elbg)
    if (allocated(relbg)) deallocate(relbg)
    if (allocated(kg)) deallocate(kg)
    call precal()
    ref_precal = errnr
    if (errnr /= 0) return
    c_lbeta = merge(1, 0, lbeta)
    c_lrandb = merge(1, 0, lrandb)
    c_sytop = sytop
    nimn = (max_nr_element_nodes*(max_nr_element_nodes+1))/2
    ntot = elanz*nimn*kwnanz
    c_elbg(1:ntot) = reshape(elbg, (/ntot/))
    if (relanz > 0) then
       c_relbg(1:relanz*nimn) = reshape(relbg, (/relanz*nimn/))
       c_kg(1:relanz*eanz*kwnanz) = reshape(kg, (/relanz*eanz*kwnanz/))
    end if
  end function ref_precal

  !> One (k,l) system exactly as assembled by the reference: kompab, then the
  !> optional Dirichlet routines, then scalab.  Returns the scaled band matrix
  !> ((mb+1)*sanz), the scaled RHS and the scaling factors.
  integer(c_int) function ref_assemble(c_l, c_k, c_sigma, c_a, c_b, c_fak) bind(C, name='ref_assemble')
    integer(c_int), value :: c_l, c_k
    complex(c_double_complex), intent(in) :: c_sigma(*)
    complex(c_double_complex), intent(out) :: c_a(*), c_b(*)
    real(c_double), intent(out) :: c_fak(*)
    complex(kind(0d0)), allocatable :: my_a(:), my_b(:)
    real(kind(0d0)), allocatable :: my_fak(:)
    call set_sigma(c_sigma)
    allocate(my_a((mb+1)*sanz), my_b(sanz), my_fak(sanz))
    call kompab(c_l, c_k, my_a, my_b)
    if (errnr /= 0) then
       ref_assemble = errnr
       return
    end if
    if (lrandb) call randb(my_a, my_b)
    if (lrandb2) call randb2(my_a, my_b)
    call scalab(my_a, my_b, my_fak)
    ref_assemble = errnr
    c_a(1:(mb+1)*sanz) = my_a
    c_b(1:sanz) = my_b
    c_fak(1:sanz) = my_fak
  end function ref_assemble

  !> In-place reference chol (src/chol.f90:7-88) on a caller band matrix.
  integer(c_int) function ref_chol(c_a) bind(C, name='ref_chol')
    complex(c_double_complex), intent(inout) :: c_a(*)
    call chol(c_a)
    ref_chol = errnr
  end function ref_chol

  !> Reference vre (src/vre.f90:5-71) with a factored band matrix.
  integer(c_int) function ref_vre(c_a, c_b, c_pot) bind(C, name='ref_vre')
    complex(c_double_complex), intent(in) :: c_a(*), c_b(*)
    complex(c_double_complex), intent(out) :: c_pot(*)
    call vre(c_a, c_b, c_pot)
    ref_vre = 0
  end function ref_vre

  subroutine set_sigma(c_sigma)
    complex(c_double_complex), intent(in) :: c_sigma(*)
    if (allocated(sigma)) deallocate(sigma)
    allocate(sigma(elanz))
    sigma = c_sigma(1:elanz)
  end subroutine set_sigma

  !> Forward potentials for all (k,l), following src/fem.f90:328-397, then
  !> rtrafo (src/rtrafo.f90) when swrtr=1.  kpot_out: sanz x eanz x kwnanz,
  !> hpot_out: sanz x eanz.  Requires ref_precal to have run.
  integer(c_int) function ref_forward(c_sigma, c_kpot, c_hpot, c_nthreads) bind(C, name='ref_forward')
    use omp_lib
    complex(c_double_complex), intent(in) :: c_sigma(*)
    complex(c_double_complex), intent(out) :: c_kpot(*), c_hpot(*)
    integer(c_int), value :: c_nthreads
    complex(kind(0d0)), allocatable :: my_a(:), my_b(:), my_pot(:)
    real(kind(0d0)), allocatable :: my_fak(:)
    integer :: k, l, j, nerr
    call set_sigma(c_sigma)
    ldc = .false.
    if (allocated(kpot)) deallocate(kpot)
    if (allocated(hpot)) deallocate(hpot)
    allocate(kpot(sanz, eanz, kwnanz), hpot(sanz, eanz))
    if (c_nthreads > 0) call omp_set_num_threads(c_nthreads)
    nerr = 0
    !$OMP PARALLEL DEFAULT(shared) PRIVATE(my_a, my_b, my_pot, my_fak, k, l, j)
    allocate(my_a((mb+1)*sanz), my_b(sanz), my_pot(sanz), my_fak(sanz))
    !$OMP DO
    do k = 1, kwnanz
       do l = 1, eanz
          if (lbeta .or. l == 1) then
             call kompab(l, k, my_a, my_b)
             if (lrandb) call randb(my_a, my_b)
             if (lrandb2) call randb2(my_a, my_b)
             call scalab(my_a, my_b, my_fak)
             call chol(my_a)
          else
             call kompb(l, my_b, my_fak)
          end if
          call vre(my_a, my_b, my_pot)
          do j = 1, sanz
             kpot(j, l, k) = my_pot(j) * dcmplx(my_fak(j))
             if (swrtr == 0) hpot(j, l) = kpot(j, l, k)
          end do
       end do
    end do
    !$OMP END DO
    deallocate(my_a, my_b, my_pot, my_fak)
    !$OMP END PARALLEL
    if (swrtr == 1) call rtrafo()
    c_kpot(1:sanz*eanz*kwnanz) = reshape(kpot, (/sanz*eanz*kwnanz/))
    c_hpot(1:sanz*eanz) = reshape(hpot, (/sanz*eanz/))
    ref_forward = errnr
  end function ref_forward

  !> The DC forward loop of src/inv.f90:343-400 (kompadc -> randdc -> scaldc
  !> -> choldc, or kompbdc when the factor is reused, then vredc and kpotdc =
  !> pot*fak), then rtrafo's DC branch when swrtr=1.  Leaves ldc set, so that
  !> ref_volt (bvolti) and ref_sens_dc (bsendc) take their DC branches.
  !> kpotdc_out: sanz x eanz x kwnanz, hpotdc_out: sanz x eanz (real).
  integer(c_int) function ref_forward_dc(c_sigma, c_kpotdc, c_hpotdc, c_nthreads) &
       bind(C, name='ref_forward_dc')
    use omp_lib
    complex(c_double_complex), intent(in) :: c_sigma(*)
    real(c_double), intent(out) :: c_kpotdc(*), c_hpotdc(*)
    integer(c_int), value :: c_nthreads
    real(kind(0d0)), allocatable :: my_a(:), my_b(:), my_fak(:)
    complex(kind(0d0)), allocatable :: my_pot(:)
    integer :: k, l, j
    call set_sigma(c_sigma)
    ldc = .true.
    if (allocated(kpotdc)) deallocate(kpotdc)
    if (allocated(hpotdc)) deallocate(hpotdc)
    allocate(kpotdc(sanz, eanz, kwnanz), hpotdc(sanz, eanz))
    if (c_nthreads > 0) call omp_set_num_threads(c_nthreads)
    !$OMP PARALLEL DEFAULT(shared) PRIVATE(my_a, my_b, my_pot, my_fak, k, l, j)
    allocate(my_a((mb+1)*sanz), my_b(sanz), my_pot(sanz), my_fak(sanz))
    !$OMP DO
    do k = 1, kwnanz
       do l = 1, eanz
          if (lbeta .or. l == 1) then
             call kompadc(l, k, my_a, my_b)
             if (lrandb) call randdc(my_a, my_b)
             if (lrandb2) call randbdc2(my_a, my_b)
             call scaldc(my_a, my_b, my_fak)
             call choldc(my_a)
          else
             call kompbdc(l, my_b, my_fak)
          end if
          call vredc(my_a, my_b, my_pot)
          do j = 1, sanz
             kpotdc(j, l, k) = dble(my_pot(j)) * my_fak(j)
             if (swrtr == 0) hpotdc(j, l) = kpotdc(j, l, k)
          end do
       end do
    end do
    !$OMP END DO
    deallocate(my_a, my_b, my_pot, my_fak)
    !$OMP END PARALLEL
    if (swrtr == 1) call rtrafo()
    c_kpotdc(1:sanz*eanz*kwnanz) = reshape(kpotdc, (/sanz*eanz*kwnanz/))
    c_hpotdc(1:sanz*eanz) = reshape(hpotdc, (/sanz*eanz/))
    ref_forward_dc = errnr
  end function ref_forward_dc

  !> One (k,l) system as the DC loop assembles it: kompadc, randdc, scaldc.
  integer(c_int) function ref_assemble_dc(c_l, c_k, c_sigma, c_a, c_b, c_fak) bind(C, name='ref_assemble_dc')
    integer(c_int), value :: c_l, c_k
    complex(c_double_complex), intent(in) :: c_sigma(*)
    real(c_double), intent(out) :: c_a(*), c_b(*), c_fak(*)
    real(kind(0d0)), allocatable :: my_a(:), my_b(:), my_fak(:)
    call set_sigma(c_sigma)
    allocate(my_a((mb+1)*sanz), my_b(sanz), my_fak(sanz))
    call kompadc(c_l, c_k, my_a, my_b)
    if (errnr /= 0) then
       ref_assemble_dc = errnr
       return
    end if
    if (lrandb) call randdc(my_a, my_b)
    if (lrandb2) call randbdc2(my_a, my_b)
    call scaldc(my_a, my_b, my_fak)
    ref_assemble_dc = errnr
    c_a(1:(mb+1)*sanz) = my_a
    c_b(1:sanz) = my_b
    c_fak(1:sanz) = my_fak
  end function ref_assemble_dc

  !> In-place reference choldc (src/choldc.f90:7-86).
  integer(c_int) function ref_choldc(c_a) bind(C, name='ref_choldc')
    real(c_double), intent(inout) :: c_a(*)
    call choldc(c_a)
    ref_choldc = errnr
  end function ref_choldc

  !> Reference bsendc (src/bsendc.f90:1-187) after ref_forward_dc and ref_volt(1).
  !> sensdc_out: nanz x elanz column major (real).
  integer(c_int) function ref_sens_dc(c_sensdc) bind(C, name='ref_sens_dc')
    real(c_double), intent(out) :: c_sensdc(*)
    if (allocated(sensdc)) deallocate(sensdc)
    allocate(sensdc(nanz, manz))
    call bsendc(.false.)
    c_sensdc(1:nanz*manz) = reshape(sensdc, (/nanz*manz/))
    ref_sens_dc = errnr
  end function ref_sens_dc

  !> Neighbours and the triangular smoothness operator: bnachbar
  !> (src/bnachbar.f90) and bsmatm with ltri = 1, i.e. bsmatmtri
  !> (src/bsmatm_mod.f90:364-461), with the element centroids as relem sets
  !> them (src/relem.f90:134-143) and c_ndec decoupled edges: element pairs
  !> c_edecoup (c_ndec x 2, column major, 1-based) with strength factors
  !> c_edecstr, the arrays rall fills from decouplings.dat (src/rall.f90:648-665).
  !> smatm_out and nachbar_out: manz x (maxnodes+1), column major.
  integer(c_int) function ref_bsmatmtri(c_alfx, c_alfz, c_ndec, c_edecoup, c_edecstr, &
       c_smatm, c_nachbar) bind(C, name='ref_bsmatmtri')
    use bsmatm_mod, only: bsmatm
    use konvmod, only: ltri, alfx, alfz
    real(c_double), value :: c_alfx, c_alfz
    integer(c_int), value :: c_ndec
    integer(c_int), intent(in) :: c_edecoup(*)
    real(c_double), intent(in) :: c_edecstr(*)
    real(c_double), intent(out) :: c_smatm(*)
    integer(c_int), intent(out) :: c_nachbar(*)
    logical :: lb
    integer :: k, iel, ti, j, ncol
    if (allocated(espx)) deallocate(espx)
    if (allocated(espy)) deallocate(espy)
    allocate(espx(elanz), espy(elanz))
    espx = 0d0
    espy = 0d0
    iel = 0
    do ti = 1, typanz
       do j = 1, nelanz(ti)
          iel = iel + 1
          if (typ(ti) < 10) then
             do k = 1, selanz(ti)
                espx(iel) = espx(iel) + sx(snr(nrel(iel, k)))
                espy(iel) = espy(iel) + sy(snr(nrel(iel, k)))
             end do
             espx(iel) = espx(iel) / selanz(ti)
             espy(iel) = espy(iel) / selanz(ti)
          end if
       end do
    end do
    if (allocated(nachbar)) deallocate(nachbar)
    call bnachbar()
    if (allocated(edecoup)) deallocate(edecoup)
    if (allocated(edecstr)) deallocate(edecstr)
    decanz = c_ndec
    allocate(edecoup(max(decanz, 1), 2), edecstr(max(decanz, 1)))
    do k = 1, decanz
       edecoup(k, 1) = c_edecoup(k)
       edecoup(k, 2) = c_edecoup(k + decanz)
       edecstr(k) = c_edecstr(k)
    end do
    ltri = 1
    alfx = c_alfx
    alfz = c_alfz
    if (allocated(smatm)) deallocate(smatm)
    call bsmatm(1, lb)
    ncol = max_nr_element_nodes + 1
    c_smatm(1:manz*ncol) = reshape(smatm, (/manz*ncol/))
    c_nachbar(1:manz*ncol) = reshape(nachbar, (/manz*ncol/))
    ref_bsmatmtri = errnr
  end function ref_bsmatmtri

  !> The rectangular smoothness operator: bsmatm with ltri = 0, i.e.
  !> bsmatmreg (src/bsmatm_mod.f90:190-362), on an nx x nz grid of
  !> quadrilateral elements numbered row by row (element (i, j) = (i-1)*nx + j).
  !> smatm_out: manz x 3, column major (diagonal, right neighbour, lower neighbour).
  integer(c_int) function ref_bsmatmreg(c_nx, c_nz, c_alfx, c_alfz, c_smatm) bind(C, name='ref_bsmatmreg')
    use bsmatm_mod, only: bsmatm
    use konvmod, only: ltri, alfx, alfz, nx, nz
    integer(c_int), value :: c_nx, c_nz
    real(c_double), value :: c_alfx, c_alfz
    real(c_double), intent(out) :: c_smatm(*)
    logical :: lb
    ltri = 0
    nx = c_nx
    nz = c_nz
    alfx = c_alfx
    alfz = c_alfz
    if (allocated(smatm)) deallocate(smatm)
    call bsmatm(1, lb)
    c_smatm(1:manz*3) = reshape(smatm(:, 1:3), (/manz*3/))
    ref_bsmatmreg = errnr
  end function ref_bsmatmreg

  !> A bounded sample of the forward loop for timing: systems s = (k-1)*eanz + l
  !> in [c_first, c_last] (1-based), each assembled, scaled, factored and
  !> solved as in src/fem.f90:355-380, OpenMP over the systems.  No output.
  integer(c_int) function ref_forward_sample(c_sigma, c_first, c_last, c_nthreads) &
       bind(C, name='ref_forward_sample')
    use omp_lib
    complex(c_double_complex), intent(in) :: c_sigma(*)
    integer(c_int), value :: c_first, c_last, c_nthreads
    complex(kind(0d0)), allocatable :: my_a(:), my_b(:), my_pot(:)
    real(kind(0d0)), allocatable :: my_fak(:)
    integer :: s, k, l
    call set_sigma(c_sigma)
    if (c_nthreads > 0) call omp_set_num_threads(c_nthreads)
    !$OMP PARALLEL DEFAULT(shared) PRIVATE(my_a, my_b, my_pot, my_fak, s, k, l)
    allocate(my_a((mb+1)*sanz), my_b(sanz), my_pot(sanz), my_fak(sanz))
    !$OMP DO SCHEDULE(dynamic, 1)
    do s = c_first, c_last
       k = (s - 1) / eanz + 1
       l = mod(s - 1, eanz) + 1
       call kompab(l, k, my_a, my_b)
       if (lrandb) call randb(my_a, my_b)
       call scalab(my_a, my_b, my_fak)
       call chol(my_a)
       call vre(my_a, my_b, my_pot)
    end do
    !$OMP END DO
    deallocate(my_a, my_b, my_pot, my_fak)
    !$OMP END PARALLEL
    ref_forward_sample = errnr
  end function ref_forward_sample

  !> Reference bvolt (CRMod, mode 0) or bvolti (CRTomo, mode 1).
  integer(c_int) function ref_volt(c_mode, c_volt, c_sigmaa) bind(C, name='ref_volt')
    integer(c_int), value :: c_mode
    complex(c_double_complex), intent(out) :: c_volt(*), c_sigmaa(*)
    kfak = 1d0
    if (c_mode == 0) then
       call bvolt()
    else
       call bvolti()
    end if
    ref_volt = errnr
    c_volt(1:nanz) = volt
    c_sigmaa(1:nanz) = sigmaa
  end function ref_volt

  !> Reference bsens (mode 0, dV/dsigma) or bsensi (mode 1, d(-lnV)/d(ln sigma)).
  !> sens_out: nanz x elanz column major.  Requires ref_forward and ref_volt.
  integer(c_int) function ref_sens(c_mode, c_sens) bind(C, name='ref_sens')
    integer(c_int), value :: c_mode
    complex(c_double_complex), intent(out) :: c_sens(*)
    real(kind(0d0)) :: c1
    if (allocated(sens)) deallocate(sens)
    allocate(sens(nanz, manz))
    if (c_mode == 0) then
       call bsens()
    else
       call bsensi(.false.)
    end if
    c_sens(1:nanz*manz) = reshape(sens, (/nanz*manz/))
    ref_sens = errnr
    c1 = 0d0
  end function ref_sens

  !> Reference chkpo2 (src/dmisft.f90:213-300) on one datum.
  integer(c_int) function ref_chkpo2(c_dati, c_sigi, c_ldum, c_cdat, c_csig, c_wdlok) &
       bind(C, name='ref_chkpo2')
    complex(c_double_complex), intent(in) :: c_dati(1), c_sigi(1)
    integer(c_int), value :: c_ldum
    complex(c_double_complex), intent(out) :: c_cdat(1), c_csig(1)
    integer(c_int), intent(out) :: c_wdlok(1)
    complex(kind(0d0)) :: cdat, csig
    integer(kind=4) :: wd
    logical :: ld
    wd = 1
    ld = c_ldum /= 0
    call chkpo2(c_dati(1), c_sigi(1), cdat, csig, wd, ld)
    c_cdat(1) = cdat
    c_csig(1) = csig
    c_wdlok(1) = wd
    ref_chkpo2 = 0
  end function ref_chkpo2

  !> Reference dmisft (src/dmisft.f90:17-208) with llam = lsetup = .false.
  !> (no text output).  Its control files eps.ctr and run.ctr are created
  !> empty in the caller's directory c_dir (at most 59 characters), which
  !> becomes ramd.  flags: 1 lpol, 2 lfpi, 4 ldiff, 8 lrobust.  c_out:
  !> nrmsd, betrms, pharms, rmssum, l1rat, npol; c_wmatd is updated as dmisft
  !> updates wmatd.
  integer(c_int) function ref_dmisft(c_nanz, c_dat, c_sigmaa, c_wmatd, c_wmatdr, c_wmatdp, &
       c_d0, c_fm0, c_flags, c_l1min, c_dir, c_dirlen, c_out) bind(C, name='ref_dmisft')
    use invmod, only: dat, wmatd, d0, fm0, lfpi
    use konvmod, only: nrmsd, betrms, pharms, rmssum, l1rat, l1min, lrobust, ldiff, llam, &
         lstep, it
    use pathmod, only: ramd, lnramd, slash
    integer(c_int), value :: c_nanz, c_flags, c_dirlen
    complex(c_double_complex), intent(in) :: c_dat(c_nanz), c_sigmaa(c_nanz), c_d0(c_nanz), &
         c_fm0(c_nanz)
    real(c_double), intent(inout) :: c_wmatd(c_nanz)
    real(c_double), intent(in) :: c_wmatdr(c_nanz), c_wmatdp(c_nanz)
    real(c_double), value :: c_l1min
    character(kind=c_char), intent(in) :: c_dir(c_dirlen)
    real(c_double), intent(out) :: c_out(6)
    integer :: i
    nanz = c_nanz
    if (allocated(dat)) deallocate(dat)
    if (allocated(sigmaa)) deallocate(sigmaa)
    if (allocated(wmatd)) deallocate(wmatd)
    if (allocated(wmatdr)) deallocate(wmatdr)
    if (allocated(wmatdp)) deallocate(wmatdp)
    if (allocated(d0)) deallocate(d0)
    if (allocated(fm0)) deallocate(fm0)
    allocate(dat(nanz), sigmaa(nanz), wmatd(nanz), wmatdr(nanz), wmatdp(nanz), d0(nanz), fm0(nanz))
    dat = c_dat; sigmaa = c_sigmaa; wmatd = c_wmatd; wmatdr = c_wmatdr; wmatdp = c_wmatdp
    d0 = c_d0; fm0 = c_fm0
    lpol = iand(c_flags, 1) /= 0
    lfpi = iand(c_flags, 2) /= 0
    ldiff = iand(c_flags, 4) /= 0
    lrobust = iand(c_flags, 8) /= 0
    l1min = c_l1min
    llam = .false.; lstep = .false.; it = 1
    ramd = ' '
    do i = 1, c_dirlen
       ramd(i:i) = c_dir(i)
    end do
    lnramd = c_dirlen
    slash = '/'
    fprun = 10; fpeps = 15
    open(fpeps, file=ramd(1:lnramd)//'/eps.ctr', status='replace')
    close(fpeps)
    open(fprun, file=ramd(1:lnramd)//'/run.ctr', status='replace')
    close(fprun)
    l1rat = 0d0
    call dmisft(.false.)
    close(fpeps)
    close(fprun)
    c_out = (/ nrmsd, betrms, pharms, rmssum, l1rat, dble(npol) /)
    c_wmatd = wmatd
    ref_dmisft = errnr
  end function ref_dmisft

  !> Reference bpar (src/bpar.f90) and bsigma (src/bsigma.f90) for the loaded
  !> mesh's elanz elements with the caller's element-to-parameter map (1-based).
  integer(c_int) function ref_bpar(c_manz, c_mnr, c_sigma, c_par) bind(C, name='ref_bpar')
    use invmod, only: par
    integer(c_int), value :: c_manz
    integer(c_int), intent(in) :: c_mnr(*)
    complex(c_double_complex), intent(in) :: c_sigma(*)
    complex(c_double_complex), intent(inout) :: c_par(c_manz)
    manz = c_manz
    if (allocated(mnr)) deallocate(mnr)
    allocate(mnr(elanz))
    mnr = c_mnr(1:elanz)
    if (allocated(sigma)) deallocate(sigma)
    allocate(sigma(elanz))
    sigma = c_sigma(1:elanz)
    if (allocated(par)) deallocate(par)
    allocate(par(manz))
    par = c_par
    call bpar()
    c_par = par
    ref_bpar = errnr
  end function ref_bpar

  integer(c_int) function ref_bsigma(c_manz, c_mnr, c_par, c_sigma) bind(C, name='ref_bsigma')
    use invmod, only: par
    integer(c_int), value :: c_manz
    integer(c_int), intent(in) :: c_mnr(*)
    complex(c_double_complex), intent(in) :: c_par(c_manz)
    complex(c_double_complex), intent(out) :: c_sigma(*)
    manz = c_manz
    if (allocated(mnr)) deallocate(mnr)
    allocate(mnr(elanz))
    mnr = c_mnr(1:elanz)
    if (allocated(sigma)) deallocate(sigma)
    allocate(sigma(elanz))
    if (allocated(par)) deallocate(par)
    allocate(par(manz))
    par = c_par
    call bsigma()
    c_sigma(1:elanz) = sigma
    ref_bsigma = errnr
  end function ref_bsigma

  !> Reference blam0 (src/blam0.f90) on the caller's J (nanz x manz, column
  !> major): mode 0 the complex branch, 1 lfpi (real parts of sens), 2 ldc
  !> (sensdc = real parts).  c_w = wmatd with wdfak = 1.  No preset lambda.
  integer(c_int) function ref_blam0(c_nanz, c_manz, c_sens, c_w, c_mode, c_alfx, c_alfz, c_lammax) &
       bind(C, name='ref_blam0')
    use invmod, only: wmatd, wdfak, lfpi
    use konvmod, only: lammax, lamnull_cri, lamnull_fpi, llamf, alfx, alfz, nz, lverb
    use femmod, only: ldc
    integer(c_int), value :: c_nanz, c_manz, c_mode
    complex(c_double_complex), intent(in) :: c_sens(c_nanz, c_manz)
    real(c_double), intent(in) :: c_w(c_nanz)
    real(c_double), value :: c_alfx, c_alfz
    real(c_double), intent(out) :: c_lammax
    nanz = c_nanz; manz = c_manz
    if (allocated(sens)) deallocate(sens)
    if (allocated(sensdc)) deallocate(sensdc)
    if (allocated(wmatd)) deallocate(wmatd)
    if (allocated(wdfak)) deallocate(wdfak)
    allocate(sens(nanz, manz), sensdc(nanz, manz), wmatd(nanz), wdfak(nanz))
    sens = c_sens
    sensdc = dble(c_sens)
    wmatd = c_w
    wdfak = 1
    ldc = c_mode == 2
    lfpi = c_mode == 1
    llamf = 0; lamnull_cri = 0d0; lamnull_fpi = 0d0; nz = 0; lverb = .false.
    alfx = c_alfx; alfz = c_alfz
    call blam0()
    c_lammax = lammax
    ref_blam0 = 0
  end function ref_blam0

  !> Reference rdati (src/rdati.f90:84-363) on the file c_path (no synthetic
  !> noise): the data vector and weights as rdati leaves them.  c_model: the
  !> error model stabw0, stabm0, stabpA1, stabpB, stabpA2, stabp0.  Outputs are
  !> sized nanz_max; c_info returns nanz, lindiv.
  integer(c_int) function ref_rdati(c_path, c_plen, c_eanz, c_ldc, c_model, c_nmax, c_info, c_strnr, &
       c_vnr, c_dat, c_wmatdr, c_wmatdp, c_wmatdcri, c_wdfak, c_lnse, c_noise) bind(C, name='ref_rdati')
    use invmod, only: dat, wmatd, wdfak, wmatd2
    use konvmod, only: lverb
    use femmod, only: ldc
    use alloci, only: rnd_r, rnd_p
    integer(c_int), value :: c_plen, c_eanz, c_ldc, c_nmax, c_lnse
    character(kind=c_char), intent(in) :: c_path(c_plen)
    real(c_double), intent(in) :: c_model(6), c_noise(7)
    integer(c_int), intent(out) :: c_info(2), c_strnr(c_nmax), c_vnr(c_nmax), c_wdfak(c_nmax)
    complex(c_double_complex), intent(out) :: c_dat(c_nmax)
    real(c_double), intent(out) :: c_wmatdr(c_nmax), c_wmatdp(c_nmax), c_wmatdcri(c_nmax)
    character(len=256) :: path
    integer :: i, kanal
    path = ' '
    do i = 1, c_plen
       path(i:i) = c_path(i)
    end do
    eanz = c_eanz
    ldc = c_ldc /= 0
    ! synthetic noise (c_noise: iseed, nstabw0, nstabm0, nstabpA1, nstabpB,
    ! nstabpA2, nstabp0); rdati writes its listing files to the working
    ! directory
    lnse = c_lnse /= 0; lverb = .false.
    iseed = nint(c_noise(1)); nstabw0 = c_noise(2); nstabm0 = c_noise(3); nstabpA1 = c_noise(4)
    nstabpB = c_noise(5); nstabpA2 = c_noise(6); nstabp0 = c_noise(7)
    if (allocated(rnd_r)) deallocate(rnd_r)
    if (allocated(rnd_p)) deallocate(rnd_p)
    stabw0 = c_model(1); stabm0 = c_model(2); stabpA1 = c_model(3); stabpB = c_model(4)
    stabpA2 = c_model(5); stabp0 = c_model(6)
    if (allocated(strnr)) deallocate(strnr)
    if (allocated(strom)) deallocate(strom)
    if (allocated(volt)) deallocate(volt)
    if (allocated(sigmaa)) deallocate(sigmaa)
    if (allocated(kfak)) deallocate(kfak)
    if (allocated(wmatdr)) deallocate(wmatdr)
    if (allocated(wmatdp)) deallocate(wmatdp)
    if (allocated(vnr)) deallocate(vnr)
    if (allocated(dat)) deallocate(dat)
    if (allocated(wmatd)) deallocate(wmatd)
    if (allocated(wmatd2)) deallocate(wmatd2)
    if (allocated(sgmaa2)) deallocate(sgmaa2)
    if (allocated(wdfak)) deallocate(wdfak)
    if (allocated(wmatd_cri)) deallocate(wmatd_cri)
    kanal = 77
    nanz = 0
    ! rdati sets lindiv only when the file's first record carries it: start
    ! from a fresh program's value
    lindiv = .false.
    call rdati(kanal, path)
    ref_rdati = errnr
    c_info(1) = nanz
    c_info(2) = merge(1, 0, lindiv)
    if (errnr /= 0 .or. nanz > c_nmax) return
    c_strnr(1:nanz) = strnr; c_vnr(1:nanz) = vnr; c_dat(1:nanz) = dat
    c_wmatdr(1:nanz) = wmatdr; c_wmatdp(1:nanz) = wmatdp; c_wmatdcri(1:nanz) = wmatd_cri
    c_wdfak(1:nanz) = wdfak
  end function ref_rdati

  !> Reference update (src/update.f90, not the lambda-search branch: llam =
  !> .false.) with the rectangular smoothness operator (ltri = 0) on the
  !> caller's J (nanz x manz, column major), weights w (wdfak = 1), data dat,
  !> model response sigmaa and model par (in/out: par + step dpar).  Runs the
  !> reference's cgfac, right-hand side, cjg -> cjggra and model update.
  !> Outputs dpar (scaled by cgfac, as update leaves it), cgfac, cgres
  !> (ncgmax+1: iterations, then the residuals) and bdpar.
  integer(c_int) function ref_update(c_nanz, c_manz, c_sens, c_w, c_dat, c_sigmaa, c_par, c_smatm, &
       c_lam, c_nx, c_nz, c_ncgmax, c_eps, c_step, c_dpar, c_cgfac, c_cgres, c_bdpar) bind(C, name='ref_update')
    use invmod, only: dat, wmatd, wdfak, par, dpar, dpar2, lfpi
    use cjgmod, only: cgres, cgres2, bvec, cgfac, eps, ncgmax, con_cjgmod, des_cjgmod
    use konvmod, only: ltri, lam, nx, nz, lverb, lw_ref, llam, step, lphi0, lprior, ldiff, bdpar
    use femmod, only: ldc
    use modelmod, only: w_ref_re, w_ref_im
    integer(c_int), value :: c_nanz, c_manz, c_nx, c_nz, c_ncgmax
    complex(c_double_complex), intent(in) :: c_sens(c_nanz, c_manz), c_dat(c_nanz), c_sigmaa(c_nanz)
    complex(c_double_complex), intent(inout) :: c_par(c_manz)
    real(c_double), intent(in) :: c_w(c_nanz), c_smatm(c_manz, 3)
    real(c_double), value :: c_lam, c_eps, c_step
    complex(c_double_complex), intent(out) :: c_dpar(c_manz)
    real(c_double), intent(out) :: c_cgfac(c_manz), c_cgres(c_ncgmax + 1), c_bdpar
    nanz = c_nanz; manz = c_manz
    if (allocated(sens)) deallocate(sens)
    if (allocated(smatm)) deallocate(smatm)
    if (allocated(wmatd)) deallocate(wmatd)
    if (allocated(wdfak)) deallocate(wdfak)
    if (allocated(dat)) deallocate(dat)
    if (allocated(sigmaa)) deallocate(sigmaa)
    if (allocated(par)) deallocate(par)
    if (allocated(dpar)) deallocate(dpar)
    if (allocated(dpar2)) deallocate(dpar2)
    if (allocated(w_ref_re)) deallocate(w_ref_re)
    if (allocated(w_ref_im)) deallocate(w_ref_im)
    allocate(sens(nanz, manz), smatm(manz, 3), wmatd(nanz), wdfak(nanz), dat(nanz), sigmaa(nanz), &
         par(manz), dpar(manz), dpar2(manz), w_ref_re(max(manz, nanz + 1)), w_ref_im(max(manz, nanz + 1)))
    sens = c_sens; smatm = c_smatm; wmatd = c_w; wdfak = 1; dat = c_dat; sigmaa = c_sigmaa
    par = c_par; dpar = dcmplx(0d0); dpar2 = dcmplx(0d0); w_ref_re = 0d0; w_ref_im = 0d0
    ltri = 0; lam = c_lam; nx = c_nx; nz = c_nz; ncgmax = c_ncgmax; eps = c_eps; step = c_step
    llam = .false.; lfpi = .false.; ldc = .false.; ldiff = .false.; lprior = .false.
    lw_ref = .false.; lphi0 = .false.; lverb = .false.
    if (allocated(cgres)) call des_cjgmod(1, fetxt, errnr)
    call con_cjgmod(1, fetxt, errnr)
    if (errnr /= 0) then
       ref_update = errnr
       return
    end if
    cgres = 0d0; cgres2 = 0d0
    errnr = 0
    call update()
    c_dpar = dpar
    c_cgfac = cgfac
    c_cgres = cgres(1:ncgmax + 1)
    c_par = par
    c_bdpar = bdpar
    ref_update = errnr
  end function ref_update

end module crt_ref_driver
