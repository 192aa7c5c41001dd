"""Host-side mirror of the CRMod/CRTomo forward + sensitivity stages over the
HIP C-ABI (include/crt_hip.h, libcrt_hip.so).

The class :class:`CRT` keeps one mesh on the device and exposes the stages the
reference calls from its serial region, under the reference's names:

=================  ==========================================  =====================
method             reference routine                           C-ABI entry point
=================  ==========================================  =====================
``rwaven``         src/rwaven.f90:22-171                       ``crt_rwaven``
``precal``         src/precal.f90:13-222                       ``crt_precal``
``forward``        src/fem.f90:328-397 (+ rtrafo)              ``crt_forward``
``bvolt``          src/bvolt.f90 / ``bvolti`` src/bvolti.f90   ``crt_voltages``
``bsens``          src/bsens.f90 / ``bsensi`` src/bsensi.f90   ``crt_sensitivities``
``adjoint_grad``   data term of src/update.f90:235-289         ``crt_adjoint_gradient``
``cg_solve``       src/update.f90:179-307 + cjggra             ``crt_cg_solve``
=================  ==========================================  =====================

Errors raise :class:`CRTError` carrying the reference's error number.  There is
no CPU fallback: without the built library or a GPU the calls fail.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

from .formats import Mesh

LIB_PATH = Path(__file__).resolve().parent / "libcrt_hip.so"

# every symbol include/crt_hip.h declares
EXPORTS = (
    "crt_create", "crt_destroy", "crt_rwaven", "crt_set_wavenumbers", "crt_precal",
    "crt_get_precal", "crt_forward", "crt_forward_dev", "crt_voltages", "crt_voltages_dev",
    "crt_sensitivities", "crt_adjoint_gradient", "crt_adjoint_gradient_dev", "crt_jhw_r",
    "crt_cg_solve", "crt_forward_dc", "crt_sensitivities_dc", "crt_set_boundary_values",
    "crt_smoothness", "crt_set_decouplings", "crt_set_graph", "crt_set_device_budget", "crt_set_tuning",
    "crt_set_recompute", "crt_set_sens_form",
    "crt_forward_many",
    "crt_error_info", "crt_set_timing", "crt_stage_times", "crt_factor_phase_ticks",
    "crt_kpot_device", "crt_dmisft", "crt_dmisft_dev", "crt_dmisft_dev_noalloc", "crt_set_model_map", "crt_bpar", "crt_bpar_dev",
    "crt_bsigma", "crt_bsigma_dev", "crt_update_step", "crt_update_step_dev", "crt_blam0",
)

# dmisft switches (the reference's lpol, lfpi, ldiff, lrobust)
MISFIT_LPOL, MISFIT_LFPI, MISFIT_LDIFF, MISFIT_LROBUST = 1, 2, 4, 8

ERRORS = {
    18: "unknown element type", 19: "band width too small for an element",
    26: "element geometry (non-positive determinant)", 27: "zero diagonal in scaling",
    28: "zero pivot in the Cholesky factorisation", 29: "electrode node out of range",
    35: "mixed-boundary geometry", 46: "electrode number out of range",
    99: "every datum excluded by its polarity",
    73: "duplicate electrodes", 82: "zero voltage", 101: "mixed and Dirichlet boundaries",
    111: "boundary element not linked to an area element",
    1000: "invalid argument", 1001: "HIP runtime failure", 1002: "stage called out of order",
    1003: "band width above the supported limit",
    1004: "size above the no-allocation entry point's limit",
}


class CRTError(RuntimeError):
    def __init__(self, where, code):
        super().__init__(f"{where}: error {code} ({ERRORS.get(code, 'unknown')})")
        self.code = code


class CMesh(C.Structure):
    _fields_ = [("sanz", C.c_int32), ("typanz", C.c_int32), ("mb", C.c_int32),
                ("eanz", C.c_int32), ("typ", C.c_void_p), ("nelanz", C.c_void_p),
                ("selanz", C.c_void_p), ("nrel", C.c_void_p), ("rnr", C.c_void_p),
                ("snr", C.c_void_p), ("sx", C.c_void_p), ("sy", C.c_void_p),
                ("enr", C.c_void_p), ("lsytop", C.c_int32), ("lsink", C.c_int32),
                ("nsink", C.c_int32)]


_lib = None


def load_library(path: Path = LIB_PATH):
    """Load libcrt_hip.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not path.exists():
        raise RuntimeError(f"{path} is missing: build it with `python -m titatest_amd.build`")
    L = C.CDLL(str(path))
    vp, i32, f64 = C.c_void_p, C.c_int32, C.c_double
    L.crt_create.argtypes = [C.POINTER(vp), C.POINTER(CMesh)]
    L.crt_destroy.argtypes = [vp]
    L.crt_rwaven.argtypes = [vp, i32, vp, vp, C.POINTER(i32), vp, vp]
    L.crt_set_wavenumbers.argtypes = [vp, i32, vp, vp, i32]
    L.crt_precal.argtypes = [vp, C.POINTER(i32), C.POINTER(i32)]
    L.crt_get_precal.argtypes = [vp, vp, vp, vp]
    L.crt_forward.argtypes = [vp, vp, vp, vp]
    L.crt_forward_dev.argtypes = [vp, vp]
    L.crt_forward_dc.argtypes = [vp, vp, vp, vp]
    L.crt_set_boundary_values.argtypes = [vp, i32, vp, vp]
    L.crt_smoothness.argtypes = [vp, i32, f64, f64, i32, i32, vp, vp]
    L.crt_set_decouplings.argtypes = [vp, i32, vp, vp]
    L.crt_set_graph.argtypes = [vp, i32]
    L.crt_set_device_budget.argtypes = [vp, C.c_uint64]
    L.crt_set_tuning.argtypes = [vp, i32, i32]
    L.crt_set_recompute.argtypes = [vp, i32]
    L.crt_set_sens_form.argtypes = [vp, i32]
    L.crt_forward_many.argtypes = [i32, vp, vp]
    L.crt_sensitivities_dc.argtypes = [vp, vp]
    L.crt_voltages.argtypes = [vp, i32, vp, vp, i32, vp, vp, vp]
    L.crt_voltages_dev.argtypes = [vp, i32, vp, vp, i32, vp, vp]
    L.crt_sensitivities.argtypes = [vp, i32, vp]
    L.crt_adjoint_gradient.argtypes = [vp, i32, vp, vp, vp]
    L.crt_adjoint_gradient_dev.argtypes = [vp, i32, vp, vp, vp]
    L.crt_jhw_r.argtypes = [vp, vp, vp, vp]
    L.crt_cg_solve.argtypes = [vp, vp, vp, vp, f64, i32, i32, vp, i32, vp, i32, vp, vp, i32, f64,
                               vp, vp, vp]
    L.crt_error_info.argtypes = [vp, C.POINTER(i32), C.POINTER(i32), C.POINTER(i32)]
    L.crt_set_timing.argtypes = [vp, i32]
    L.crt_stage_times.argtypes = [vp, i32, vp, vp, vp]
    L.crt_kpot_device.argtypes = [vp, C.POINTER(vp)]
    L.crt_factor_phase_ticks.argtypes = [vp, i32, vp]
    L.crt_dmisft.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, vp, i32, f64, vp, vp]
    L.crt_dmisft_dev.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, vp, i32, f64, vp, vp]
    L.crt_dmisft_dev_noalloc.argtypes = [vp, i32, vp, vp, vp, vp, vp, vp, vp, i32, f64, vp, vp]
    L.crt_set_model_map.argtypes = [vp, i32, vp]
    L.crt_bpar.argtypes = [vp, vp, vp]
    L.crt_bpar_dev.argtypes = [vp, vp, vp]
    L.crt_bsigma.argtypes = [vp, vp, vp]
    L.crt_bsigma_dev.argtypes = [vp, vp, vp]
    L.crt_update_step.argtypes = [vp, i32, vp, vp, f64, i32, vp]
    L.crt_update_step_dev.argtypes = [vp, i32, vp, vp, f64, i32, vp]
    L.crt_blam0.argtypes = [vp, vp, i32, f64, f64, C.POINTER(f64)]
    _lib = L
    return L


def _ptr(a):
    return None if a is None else a.ctypes.data


def forward_many(handles, sigma_dev_ptrs):
    """crt_forward_many over CRT objects, with each model's conductivities
    given as a device pointer (e.g. a torch tensor's data_ptr())."""
    L = load_library()
    n = len(handles)
    hs = (C.c_void_p * n)(*[g.h.value for g in handles])
    ps = (C.c_void_p * n)(*[int(p) for p in sigma_dev_ptrs])
    rc = L.crt_forward_many(n, hs, ps)
    if rc:
        raise CRTError("forward_many", rc)


def _c128(a):
    return np.ascontiguousarray(a, np.complex128)


class CRT:
    """One mesh on the GPU; stages mirror the reference's call sequence."""

    def __init__(self, mesh: Mesh, lsytop=False, lsink=False, nsink=0):
        self.lib = load_library()
        self.mesh = mesh
        self._keep = dict(
            typ=np.ascontiguousarray(mesh.typ, np.int32),
            nelanz=np.ascontiguousarray(mesh.nelanz, np.int32),
            selanz=np.ascontiguousarray(mesh.selanz, np.int32),
            nrel=np.asfortranarray(mesh.nrel, np.int32).ravel(order="F").copy(),
            rnr=np.ascontiguousarray(mesh.rnr if len(mesh.rnr) else np.zeros(1), np.int32),
            snr=np.ascontiguousarray(mesh.snr, np.int32),
            sx=np.ascontiguousarray(mesh.sx, np.float64),
            sy=np.ascontiguousarray(mesh.sy, np.float64),
            enr=np.ascontiguousarray(mesh.enr, np.int32),
        )
        k = self._keep
        cm = CMesh(mesh.sanz, mesh.typanz, mesh.mb, mesh.eanz,
                   *(k[n].ctypes.data for n in ("typ", "nelanz", "selanz", "nrel", "rnr", "snr",
                                                "sx", "sy", "enr")),
                   int(lsytop), int(lsink), int(nsink))
        h = C.c_void_p()
        self._check("crt_create", self.lib.crt_create(C.byref(h), C.byref(cm)))
        self.h = h
        self.elanz, self.relanz = mesh.elanz, mesh.relanz
        self.nimn = mesh.maxnodes * (mesh.maxnodes + 1) // 2
        self.K = 0
        self.nanz = 0

    def close(self):
        if getattr(self, "h", None):
            self.lib.crt_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _check(self, where, rc):
        if rc != 0:
            raise CRTError(where, rc)

    # A1 -------------------------------------------------------------------
    def rwaven(self, strnr, vnr):
        kwn = np.zeros(30)
        kwnwi = np.zeros(30)
        K = C.c_int32()
        st = np.ascontiguousarray(strnr, np.int32)
        vn = np.ascontiguousarray(vnr, np.int32)
        self._check("rwaven", self.lib.crt_rwaven(self.h, len(st), _ptr(st), _ptr(vn), C.byref(K),
                                                  _ptr(kwn), _ptr(kwnwi)))
        self.K = K.value
        self.kwn, self.kwnwi = kwn[:self.K].copy(), kwnwi[:self.K].copy()
        return self.kwn, self.kwnwi

    def set_wavenumbers(self, kwn, kwnwi, swrtr=1):
        kwn = np.ascontiguousarray(kwn, np.float64)
        kwnwi = np.ascontiguousarray(kwnwi, np.float64)
        self._check("set_wavenumbers", self.lib.crt_set_wavenumbers(self.h, len(kwn), _ptr(kwn),
                                                                    _ptr(kwnwi), swrtr))
        self.K = len(kwn)
        self.kwn, self.kwnwi = kwn, kwnwi

    # A2/A3 ----------------------------------------------------------------
    def precal(self):
        lb, lr = C.c_int32(), C.c_int32()
        self._check("precal", self.lib.crt_precal(self.h, C.byref(lb), C.byref(lr)))
        self.lbeta, self.lrandb = lb.value, lr.value
        return self.lbeta, self.lrandb

    def get_precal(self):
        K, E = self.K, self.mesh.eanz
        elbg = np.zeros(self.elanz * self.nimn * K)
        relbg = np.zeros(max(self.relanz, 1) * self.nimn)
        kg = np.zeros(max(self.relanz, 1) * E * K)
        self._check("get_precal", self.lib.crt_get_precal(self.h, _ptr(elbg), _ptr(relbg), _ptr(kg)))
        return elbg, relbg, kg

    def set_boundary_values(self, rwdnr, rwd):
        """Prescribed boundary potentials for randb2 (rrandb's rwdnr/rwd);
        empty arrays switch them off."""
        rn = np.ascontiguousarray(rwdnr, np.int32)
        rv = np.ascontiguousarray(rwd, np.complex128)
        self._check("set_boundary_values",
                    self.lib.crt_set_boundary_values(self.h, len(rn), _ptr(rn) if len(rn) else None,
                                                     _ptr(rv) if len(rv) else None))

    # A4..A11 --------------------------------------------------------------
    def forward(self, sigma, want_kpot=False, want_hpot=False):
        n, E, K = self.mesh.sanz, self.mesh.eanz, self.K
        sig = _c128(sigma)
        kpot = np.zeros(n * E * K, np.complex128) if want_kpot else None
        hpot = np.zeros(n * E, np.complex128) if want_hpot else None
        self._check("forward", self.lib.crt_forward(self.h, _ptr(sig), _ptr(kpot), _ptr(hpot)))
        out = []
        if want_kpot:
            out.append(kpot.reshape((n, E, K), order="F"))
        if want_hpot:
            out.append(hpot.reshape((n, E), order="F"))
        return tuple(out)

    # A19: DC twins ----------------------------------------------------------
    def forward_dc(self, sigma, want_kpot=False, want_hpot=False):
        """kompadc ... vredc on the device (crt_forward_dc): real kpotdc
        (sanz, eanz, K) and hpotdc (sanz, eanz) as requested.  bvolti then
        gives the DC voltages and choldc's error rules apply (27, 28)."""
        n, E, K = self.mesh.sanz, self.mesh.eanz, self.K
        sig = _c128(sigma)
        kpot = np.zeros(n * E * K) if want_kpot else None
        hpot = np.zeros(n * E) if want_hpot else None
        self._check("forward_dc", self.lib.crt_forward_dc(self.h, _ptr(sig), _ptr(kpot), _ptr(hpot)))
        out = []
        if want_kpot:
            out.append(kpot.reshape((n, E, K), order="F"))
        if want_hpot:
            out.append(hpot.reshape((n, E), order="F"))
        return tuple(out)

    def bsendc(self):
        """Real DC sensitivities (nanz, elanz) after forward_dc and bvolti."""
        out = np.zeros(self.nanz * self.elanz)
        self._check("sensitivities_dc", self.lib.crt_sensitivities_dc(self.h, _ptr(out)))
        return out.reshape((self.nanz, self.elanz), order="F")

    # A12 ------------------------------------------------------------------
    def voltages(self, strnr, vnr, mode=0, kfak=None):
        st = np.ascontiguousarray(strnr, np.int32)
        vn = np.ascontiguousarray(vnr, np.int32)
        N = len(st)
        volt = np.zeros(N, np.complex128)
        sigmaa = np.zeros(N, np.complex128)
        kf = None if kfak is None else np.ascontiguousarray(kfak, np.float64)
        self._check("voltages", self.lib.crt_voltages(self.h, N, _ptr(st), _ptr(vn), mode, _ptr(kf),
                                                      _ptr(volt), _ptr(sigmaa)))
        self.nanz = N
        return volt, sigmaa

    def bvolt(self, strnr, vnr, kfak=None):
        return self.voltages(strnr, vnr, 0, kfak)

    def bvolti(self, strnr, vnr):
        return self.voltages(strnr, vnr, 1)

    # A14 ------------------------------------------------------------------
    def sensitivities(self, mode=0, to_host=True):
        out = np.zeros(self.nanz * self.elanz, np.complex128) if to_host else None
        self._check("sensitivities", self.lib.crt_sensitivities(self.h, mode, _ptr(out)))
        return None if out is None else out.reshape((self.nanz, self.elanz), order="F")

    def bsens(self):
        return self.sensitivities(0)

    def bsensi(self):
        return self.sensitivities(1)

    # adjoint gradient ---------------------------------------------------------
    def adjoint_grad(self, w, resid, mode=1):
        w = np.ascontiguousarray(w, np.float64)
        r = _c128(resid)
        g = np.zeros(self.elanz, np.complex128)
        self._check("adjoint_gradient", self.lib.crt_adjoint_gradient(self.h, mode, _ptr(w), _ptr(r),
                                                                      _ptr(g)))
        return g

    def jhw_r(self, w, resid):
        """J^H W r from the explicit Jacobian of the last sensitivities call."""
        w = np.ascontiguousarray(w, np.float64)
        r = _c128(resid)
        g = np.zeros(self.elanz, np.complex128)
        self._check("jhw_r", self.lib.crt_jhw_r(self.h, _ptr(w), _ptr(r), _ptr(g)))
        return g

    # A16/A17 ----------------------------------------------------------------
    def set_decouplings(self, edecoup, edecstr):
        """Decoupled edges of the triangular operator (decouplings.dat):
        edecoup (decanz, 2) 1-based element pairs, edecstr (decanz,) factors."""
        ed = np.asfortranarray(edecoup, np.int32).reshape(-1, 2).ravel(order="F").copy()
        es = np.ascontiguousarray(edecstr, np.float64)
        self._check("set_decouplings", self.lib.crt_set_decouplings(self.h, len(es), _ptr(ed), _ptr(es)))

    def smoothness(self, ltri, alfx, alfz, nx=0, nz=0):
        """bsmatm's smoothness operator: smatm (manz, 3) for ltri 0, or
        (smatm, nachbar), both (manz, maxnodes+1), for ltri 1/2."""
        M, mx = self.elanz, self.mesh.maxnodes
        ncol = 3 if ltri == 0 else mx + 1
        sm = np.zeros(M * ncol)
        nb = None if ltri == 0 else np.zeros(M * (mx + 1), np.int32)
        self._check("smoothness", self.lib.crt_smoothness(self.h, ltri, float(alfx), float(alfz),
                                                          nx, nz, _ptr(sm), _ptr(nb)))
        sm = sm.reshape((M, ncol), order="F")
        return sm if ltri == 0 else (sm, nb.reshape((M, mx + 1), order="F"))

    def set_tuning(self, factor_threads=0, backsolve_depth=0):
        """Factor workgroup threads (0, 256, 512) and back-substitution
        prefetch depth (0, 8, 16) for this handle's forwards."""
        self._check("set_tuning", self.lib.crt_set_tuning(self.h, int(factor_threads), int(backsolve_depth)))

    def set_sens_form(self, form):
        """Sensitivity kernel form: 0 built-in rule, 1 pair form, 2 Gram
        form on the matrix cores (crt_set_sens_form)."""
        self._check("set_sens_form", self.lib.crt_set_sens_form(self.h, int(form)))

    def set_recompute(self, wavenumbers_per_pass):
        """Hold at most this many wavenumbers' potentials at once (0 = store
        them all): forwards keep the electrode rows only, the adjoint
        gradient re-solves pass by pass (crt_set_recompute)."""
        self._check("set_recompute", self.lib.crt_set_recompute(self.h, int(wavenumbers_per_pass)))

    def set_device_budget(self, nbytes):
        """Cap the device memory a forward keeps for factors in flight."""
        self._check("set_device_budget", self.lib.crt_set_device_budget(self.h, int(nbytes)))

    # inversion-loop pieces -----------------------------------------------
    def dmisft(self, dat, sigmaa, wmatd, wmatdr, wmatdp, flags=0, l1min=1.0, d0=None, fm0=None):
        """Data misfit of src/dmisft.f90: returns (stats, wmatd_out, wdlok),
        stats = nrmsd, betrms, pharms, rmssum, l1rat, npol."""
        N = len(dat)
        w = np.array(wmatd, np.float64)
        st = np.zeros(6)
        wd = np.zeros(N, np.int32)
        d0 = None if d0 is None else _c128(d0)
        fm0 = None if fm0 is None else _c128(fm0)
        self._check("dmisft", self.lib.crt_dmisft(
            self.h, N, _ptr(_c128(dat)), _ptr(_c128(sigmaa)), _ptr(w),
            _ptr(np.ascontiguousarray(wmatdr, np.float64)), _ptr(np.ascontiguousarray(wmatdp, np.float64)),
            _ptr(d0), _ptr(fm0), int(flags), float(l1min), _ptr(st), _ptr(wd)))
        return st, w, wd

    def set_model_map(self, manz=0, mnr=None):
        """Element-to-parameter map (1-based, elanz entries), None = identity."""
        self._mnr = None if mnr is None else np.ascontiguousarray(mnr, np.int32)
        self._manz = self.elanz if mnr is None else int(manz)
        self._check("set_model_map", self.lib.crt_set_model_map(self.h, self._manz, _ptr(self._mnr)))

    def bpar(self, sigma, par=None):
        manz = getattr(self, "_manz", self.elanz)
        out = np.zeros(manz, np.complex128) if par is None else np.array(par, np.complex128)
        self._check("bpar", self.lib.crt_bpar(self.h, _ptr(_c128(sigma)), _ptr(out)))
        return out

    def bsigma(self, par):
        out = np.zeros(self.elanz, np.complex128)
        self._check("bsigma", self.lib.crt_bsigma(self.h, _ptr(_c128(par)), _ptr(out)))
        return out

    def update_step(self, par, dpar, step, lphi0=False):
        """par + step dpar (src/update.f90:332-351): (par_new, bdpar, forced)."""
        p = np.array(par, np.complex128)
        out = np.zeros(2)
        self._check("update_step", self.lib.crt_update_step(self.h, len(p), _ptr(p), _ptr(_c128(dpar)),
                                                            float(step), int(bool(lphi0)), _ptr(out)))
        return p, out[0], int(out[1])

    def blam0(self, w, mode=0, alfx=1.0, alfz=1.0):
        """Starting lambda of src/blam0.f90 from the explicit J on the device."""
        out = C.c_double()
        self._check("blam0", self.lib.crt_blam0(self.h, _ptr(np.ascontiguousarray(w, np.float64)), int(mode),
                                                float(alfx), float(alfz), C.byref(out)))
        return out.value

    def set_graph(self, on=True):
        """Run crt_cg_solve's iterations as one captured HIP graph."""
        self._check("set_graph", self.lib.crt_set_graph(self.h, int(bool(on))))

    def cg_solve(self, w, dat, sigmaa, lam, ltri, smatm, rdiag, ncgmax, eps=1e-4, nx=0,
                 nachbar=None, rm=None):
        M = self.elanz
        w = np.ascontiguousarray(w, np.float64)
        dat = _c128(dat)
        sigmaa = _c128(sigmaa)
        sm = np.asfortranarray(np.atleast_2d(np.asarray(smatm, np.float64).reshape(M, -1)))
        smf = sm.ravel(order="F").copy()
        nb = None
        maxnodes = 0
        if nachbar is not None:
            nbf = np.asfortranarray(np.asarray(nachbar, np.int32))
            maxnodes = nbf.shape[1]
            nb = nbf.ravel(order="F").copy()
        rd = np.ascontiguousarray(rdiag, np.float64)
        rmv = None if rm is None else _c128(rm)
        dpar = np.zeros(M, np.complex128)
        cgfac = np.zeros(M)
        cgres = np.zeros(ncgmax + 2)
        self._check("cg_solve", self.lib.crt_cg_solve(
            self.h, _ptr(w), _ptr(dat), _ptr(sigmaa), float(lam), ltri, nx, _ptr(smf), sm.shape[1],
            _ptr(nb), maxnodes, _ptr(rd), _ptr(rmv), ncgmax, float(eps), _ptr(dpar), _ptr(cgfac),
            _ptr(cgres)))
        return dpar, cgfac, cgres

    # timing -----------------------------------------------------------------
    def set_timing(self, on=True):
        self.lib.crt_set_timing(self.h, 1 if on else 0)

    def stage_times(self):
        nmax = 64
        names = C.create_string_buffer(48 * nmax)
        ms = np.zeros(nmax)
        cnt = np.zeros(nmax, np.int32)
        k = self.lib.crt_stage_times(self.h, nmax, names, _ptr(ms), _ptr(cnt))
        raw = names.raw
        out = {}
        for i in range(k):
            nm = raw[48 * i:48 * (i + 1)].split(b"\0", 1)[0].decode()
            out[nm] = (float(ms[i]), int(cnt[i]))
        return out

    # wave 0 (panel wave) then waves 1.. (trailing waves), per stage phase
    PHASES = ("w0_prologue", "w0_wait_lookahead", "w0_panel", "w0_wait_trailing",
              "w1_prologue_lookahead", "w1_wait_lookahead", "w1_trailing", "w1_wait_panel")

    def factor_phases(self, enable=True):
        """Return the factor kernel's phase times (ms summed over workgroups)
        accumulated since the previous call, and switch the clock on/off."""
        t = np.zeros(8, np.uint64)
        self._check("factor_phase_ticks", self.lib.crt_factor_phase_ticks(self.h, int(enable), _ptr(t)))
        return {k: float(v) * 1e-5 for k, v in zip(self.PHASES, t)}

    def error_info(self):
        c, k, l = C.c_int32(), C.c_int32(), C.c_int32()
        self.lib.crt_error_info(self.h, C.byref(c), C.byref(k), C.byref(l))
        return c.value, k.value, l.value
