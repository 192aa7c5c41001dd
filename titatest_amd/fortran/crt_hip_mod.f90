!> iso_c_binding interface to libcrt_hip.so (include/crt_hip.h).
!>
!> This is the module a CRMod/CRTomo host `USE`s to hand its stages to the
!> MI355X: the arrays it passes are its own module arrays (elemmod, electrmod,
!> sigmamod, datmod, alloci), unchanged, column major with 1-based index values.
!> Complex arrays are COMPLEX(KIND(0D0)) and travel as C double pointers.
module crt_hip
  use iso_c_binding
  implicit none

  !> Mirrors struct crt_mesh.
  type, bind(C) :: crt_mesh
     integer(c_int32_t) :: sanz, typanz, mb, eanz
     type(c_ptr) :: typ, nelanz, selanz, nrel, rnr, snr, sx, sy, enr
     integer(c_int32_t) :: lsytop, lsink, nsink
  end type crt_mesh

  interface
     integer(c_int) function crt_create(h, mesh) bind(C, name='crt_create')
       import :: c_int, c_ptr, crt_mesh
       type(c_ptr) :: h
       type(crt_mesh), intent(in) :: mesh
     end function crt_create

     integer(c_int) function crt_destroy(h) bind(C, name='crt_destroy')
       import :: c_int, c_ptr
       type(c_ptr), value :: h
     end function crt_destroy

     integer(c_int) function crt_rwaven(h, nanz, strnr, vnr, kwnanz, kwn, kwnwi) &
          bind(C, name='crt_rwaven')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nanz
       integer(c_int32_t), intent(in) :: strnr(*), vnr(*)
       integer(c_int32_t), intent(out) :: kwnanz
       real(c_double), intent(out) :: kwn(30), kwnwi(30)
     end function crt_rwaven

     integer(c_int) function crt_set_wavenumbers(h, kwnanz, kwn, kwnwi, swrtr) &
          bind(C, name='crt_set_wavenumbers')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: kwnanz, swrtr
       real(c_double), intent(in) :: kwn(*), kwnwi(*)
     end function crt_set_wavenumbers

     integer(c_int) function crt_precal(h, lbeta, lrandb) bind(C, name='crt_precal')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: h
       integer(c_int32_t), intent(out) :: lbeta, lrandb
     end function crt_precal

     integer(c_int) function crt_forward(h, sigma, kpot, hpot) bind(C, name='crt_forward')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex), intent(in) :: sigma(*)
       type(c_ptr), value :: kpot, hpot
     end function crt_forward

     integer(c_int) function crt_set_boundary_values(h, rwdanz, rwdnr, rwd) &
          bind(C, name='crt_set_boundary_values')
       import :: c_int, c_ptr, c_int32_t, c_double_complex
       type(c_ptr), value :: h
       integer(c_int32_t), value :: rwdanz
       integer(c_int32_t), intent(in) :: rwdnr(*)
       complex(c_double_complex), intent(in) :: rwd(*)
     end function crt_set_boundary_values

     integer(c_int) function crt_smoothness(h, ltri, alfx, alfz, nx, nz, smatm, nachbar) &
          bind(C, name='crt_smoothness')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: ltri, nx, nz
       real(c_double), value :: alfx, alfz
       real(c_double), intent(out) :: smatm(*)
       type(c_ptr), value :: nachbar
     end function crt_smoothness

     ! decoupled edges (decouplings.dat): edecoup(decanz, 2), edecstr(decanz)
     integer(c_int) function crt_set_decouplings(h, decanz, edecoup, edecstr) &
          bind(C, name='crt_set_decouplings')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       integer(c_int32_t), value :: decanz
       integer(c_int32_t), intent(in) :: edecoup(*)
       real(c_double), intent(in) :: edecstr(*)
     end function crt_set_decouplings

     ! DC twins (ldc): kpotdc/hpotdc real, or c_null_ptr
     integer(c_int) function crt_forward_dc(h, sigma, kpotdc, hpotdc) bind(C, name='crt_forward_dc')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex), intent(in) :: sigma(*)
       type(c_ptr), value :: kpotdc, hpotdc
     end function crt_forward_dc

     integer(c_int) function crt_voltages(h, nanz, strnr, vnr, mode, kfak, volt, sigmaa) &
          bind(C, name='crt_voltages')
       import :: c_int, c_ptr, c_int32_t, c_double_complex
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nanz, mode
       integer(c_int32_t), intent(in) :: strnr(*), vnr(*)
       type(c_ptr), value :: kfak
       complex(c_double_complex), intent(out) :: volt(*), sigmaa(*)
     end function crt_voltages

     integer(c_int) function crt_sensitivities(h, mode, sens) bind(C, name='crt_sensitivities')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: h
       integer(c_int32_t), value :: mode
       type(c_ptr), value :: sens
     end function crt_sensitivities

     integer(c_int) function crt_sensitivities_dc(h, sensdc) bind(C, name='crt_sensitivities_dc')
       import :: c_int, c_ptr, c_double
       type(c_ptr), value :: h
       real(c_double), intent(out) :: sensdc(*)
     end function crt_sensitivities_dc

     integer(c_int) function crt_adjoint_gradient(h, mode, w, resid, grad) &
          bind(C, name='crt_adjoint_gradient')
       import :: c_int, c_ptr, c_int32_t, c_double, c_double_complex
       type(c_ptr), value :: h
       integer(c_int32_t), value :: mode
       real(c_double), intent(in) :: w(*)
       complex(c_double_complex), intent(in) :: resid(*)
       complex(c_double_complex), intent(out) :: grad(*)
     end function crt_adjoint_gradient

     integer(c_int) function crt_cg_solve(h, w, dat, sigmaa, lam, ltri, nx, smatm, smatm_cols, &
          nachbar, maxnodes, rdiag, rm, ncgmax, eps, dpar, cgfac, cgres) bind(C, name='crt_cg_solve')
       import :: c_int, c_ptr, c_int32_t, c_double, c_double_complex
       type(c_ptr), value :: h
       real(c_double), intent(in) :: w(*)
       complex(c_double_complex), intent(in) :: dat(*), sigmaa(*)
       real(c_double), value :: lam, eps
       integer(c_int32_t), value :: ltri, nx, smatm_cols, maxnodes, ncgmax
       real(c_double), intent(in) :: smatm(*), rdiag(*)
       type(c_ptr), value :: nachbar, rm
       complex(c_double_complex), intent(out) :: dpar(*)
       real(c_double), intent(out) :: cgfac(*), cgres(*)
     end function crt_cg_solve

     ! inversion-loop pieces (src/dmisft.f90, src/bpar.f90, src/bsigma.f90,
     ! src/update.f90:332-351); d0/fm0 and wdlok may be c_null_ptr
     integer(c_int) function crt_dmisft(h, nanz, dat, sigmaa, wmatd, wmatdr, wmatdp, d0, fm0, flags, &
          l1min, stats, wdlok) bind(C, name='crt_dmisft')
       import :: c_int, c_ptr, c_int32_t, c_double, c_double_complex
       type(c_ptr), value :: h
       integer(c_int32_t), value :: nanz, flags
       complex(c_double_complex), intent(in) :: dat(*), sigmaa(*)
       real(c_double), intent(inout) :: wmatd(*)
       real(c_double), intent(in) :: wmatdr(*), wmatdp(*)
       type(c_ptr), value :: d0, fm0, wdlok
       real(c_double), value :: l1min
       real(c_double), intent(out) :: stats(6)
     end function crt_dmisft

     integer(c_int) function crt_set_model_map(h, manz, mnr) bind(C, name='crt_set_model_map')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: h
       integer(c_int32_t), value :: manz
       integer(c_int32_t), intent(in) :: mnr(*)
     end function crt_set_model_map

     integer(c_int) function crt_bpar(h, sigma, par) bind(C, name='crt_bpar')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex), intent(in) :: sigma(*)
       complex(c_double_complex), intent(inout) :: par(*)
     end function crt_bpar

     integer(c_int) function crt_bsigma(h, par, sigma) bind(C, name='crt_bsigma')
       import :: c_int, c_ptr, c_double_complex
       type(c_ptr), value :: h
       complex(c_double_complex), intent(in) :: par(*)
       complex(c_double_complex), intent(out) :: sigma(*)
     end function crt_bsigma

     integer(c_int) function crt_update_step(h, manz, par, dpar, step, lphi0, out) &
          bind(C, name='crt_update_step')
       import :: c_int, c_ptr, c_int32_t, c_double, c_double_complex
       type(c_ptr), value :: h
       integer(c_int32_t), value :: manz, lphi0
       complex(c_double_complex), intent(inout) :: par(*)
       complex(c_double_complex), intent(in) :: dpar(*)
       real(c_double), value :: step
       real(c_double), intent(out) :: out(2)
     end function crt_update_step

     integer(c_int) function crt_blam0(h, w, mode, alfx, alfz, lammax) bind(C, name='crt_blam0')
       import :: c_int, c_ptr, c_int32_t, c_double
       type(c_ptr), value :: h
       real(c_double), intent(in) :: w(*)
       integer(c_int32_t), value :: mode
       real(c_double), value :: alfx, alfz
       real(c_double), intent(out) :: lammax
     end function crt_blam0

     integer(c_int) function crt_error_info(h, code, k, l) bind(C, name='crt_error_info')
       import :: c_int, c_ptr, c_int32_t
       type(c_ptr), value :: h
       integer(c_int32_t), intent(out) :: code, k, l
     end function crt_error_info
  end interface
end module crt_hip
