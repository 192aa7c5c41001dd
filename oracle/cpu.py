"""ctypes binding of the C oracle (oracle/crt_oracle.c).  Test infrastructure only."""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "liboracle.so"

_i = np.ctypeslib.ndpointer(np.int32, flags="C")
_d = np.ctypeslib.ndpointer(np.float64, flags="C")
_z = np.ctypeslib.ndpointer(np.complex128, flags="C")


class OrcMesh(C.Structure):
    _fields_ = [("sanz", C.c_int), ("typanz", C.c_int), ("mb", C.c_int), ("eanz", C.c_int),
                ("typ", C.c_void_p), ("nelanz", C.c_void_p), ("selanz", C.c_void_p),
                ("nrel", C.c_void_p), ("rnr", C.c_void_p), ("snr", C.c_void_p),
                ("sx", C.c_void_p), ("sy", C.c_void_p), ("enr", C.c_void_p),
                ("lsytop", C.c_int), ("lsink", C.c_int), ("nsink", C.c_int),
                ("rwdanz", C.c_int), ("rwdnr", C.c_void_p), ("rwd", C.c_void_p)]


def build(force=False):
    if force or not LIB.exists() or LIB.stat().st_mtime < (HERE / "crt_oracle.c").stat().st_mtime:
        subprocess.run(["make", "-C", str(HERE), "liboracle.so"], check=True,
                       stdout=subprocess.DEVNULL)


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            build()
        L = C.CDLL(str(LIB))
        P = C.POINTER(OrcMesh)
        L.orc_rwaven.argtypes = [P, C.c_int, _i, _i, C.POINTER(C.c_int), _d, _d,
                                 C.POINTER(C.c_double), C.POINTER(C.c_double)]
        L.orc_precal.argtypes = [P, C.c_int, _d, _d, _d, _d, C.POINTER(C.c_int),
                                 C.POINTER(C.c_int), C.POINTER(C.c_double)]
        L.orc_assemble.argtypes = [P, C.c_int, _d, _d, _d, C.c_int, _z, C.c_int, C.c_int,
                                   _z, _z, _d]
        L.orc_chol.argtypes = [C.c_int, C.c_int, _z]
        L.orc_vre.argtypes = [C.c_int, C.c_int, _z, _z, _z]
        L.orc_vre.restype = None
        L.orc_solve_ld.argtypes = [C.c_int, C.c_int, _z, _z, _z]
        L.orc_forward.argtypes = [P, C.c_int, _d, C.c_int, _d, _d, _d, C.c_int, C.c_int, _z,
                                  _z, _z, C.c_int, C.c_int, C.c_int]
        L.orc_volt.argtypes = [P, C.c_int, _i, _i, _z, C.c_int, C.c_void_p, _z, _z]
        L.orc_sens.argtypes = [P, C.c_int, _i, _i, C.c_int, _d, C.c_int, _d, _z, C.c_int, _z,
                               _z, _z, C.c_int]
        L.orc_update_rhs.argtypes = [C.c_int, C.c_int, _z, _d, _i, _z, _z, C.c_double, C.c_int,
                                     _d, C.c_void_p, _d, _z]
        L.orc_update_rhs.restype = None
        L.orc_cjggra.argtypes = [C.c_int, C.c_int, _z, _d, _i, _d, _z, C.c_double, C.c_int,
                                 C.c_int, _d, C.c_int, C.c_void_p, C.c_int, C.c_int, C.c_double,
                                 _z, _d]
        L.orc_bsmatmreg.argtypes = [P, C.c_int, C.c_int, C.c_double, C.c_double, _d]
        L.orc_bsmatmreg.restype = None
        L.orc_rm_reg.argtypes = [C.c_int, C.c_int, _d, _z, _z]
        L.orc_bnachbar.argtypes = [P, _i]
        L.orc_bnachbar.restype = None
        L.orc_bsmatmtri.argtypes = [P, C.c_double, C.c_double, _i, C.c_int, _i, _d, _d]
        L.orc_bsmatmtri.restype = None
        L.orc_rm_reg.restype = None
        L.orc_assemble_dc.argtypes = [P, _d, _d, _d, C.c_int, _z, C.c_int, C.c_int, _d, _d, _d]
        L.orc_choldc.argtypes = [C.c_int, C.c_int, _d]
        L.orc_vredc.argtypes = [C.c_int, C.c_int, _d, _d, _d]
        L.orc_vredc.restype = None
        L.orc_forward_dc.argtypes = [P, C.c_int, _d, C.c_int, _d, _d, _d, C.c_int, C.c_int, _z,
                                     _d, _d, C.c_int]
        L.orc_sens_dc.argtypes = [P, C.c_int, _i, _i, C.c_int, _d, C.c_int, _d, _d, _z, _z, _d,
                                  C.c_int]
        L.orc_adjoint_grad.argtypes = [C.c_int, C.c_int, _z, _d, _i, _z, _z]
        L.orc_adjoint_grad.restype = None
        _lib = L
    return _lib


class OracleError(RuntimeError):
    def __init__(self, where, code):
        super().__init__(f"{where}: reference error code {code}")
        self.code = code


class Oracle:
    """Oracle state for one mesh: mirrors the reference's module arrays."""

    def __init__(self, mesh, lsytop=False, lsink=False, nsink=0):
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
        self.cm = OrcMesh(mesh.sanz, mesh.typanz, mesh.mb, mesh.eanz,
                          *(k[n].ctypes.data for n in ("typ", "nelanz", "selanz", "nrel", "rnr",
                                                       "snr", "sx", "sy", "enr")),
                          int(lsytop), int(lsink), int(nsink), 0, None, None)
        self.elanz = mesh.elanz
        self.relanz = mesh.relanz
        self.nimn = mesh.maxnodes * (mesh.maxnodes + 1) // 2
        self.swrtr = 1

    def set_boundary_values(self, rwdnr, rwd):
        """Prescribed boundary values (rrandb's rwdnr/rwd, applied by randb2 /
        randbdc2 after randb); empty arrays switch them off."""
        self._keep["rwdnr"] = np.ascontiguousarray(rwdnr, np.int32)
        self._keep["rwd"] = np.ascontiguousarray(rwd, np.complex128)
        self.cm.rwdanz = len(self._keep["rwdnr"])
        self.cm.rwdnr = self._keep["rwdnr"].ctypes.data if self.cm.rwdanz else None
        self.cm.rwd = self._keep["rwd"].ctypes.data if self.cm.rwdanz else None

    # --- A1 -----------------------------------------------------------------
    def rwaven(self, strnr, vnr):
        K = C.c_int()
        kwn = np.zeros(30)
        kwnwi = np.zeros(30)
        amin = C.c_double()
        amax = C.c_double()
        e = lib().orc_rwaven(C.byref(self.cm), len(strnr), np.ascontiguousarray(strnr, np.int32),
                             np.ascontiguousarray(vnr, np.int32), C.byref(K), kwn, kwnwi,
                             C.byref(amin), C.byref(amax))
        if e:
            raise OracleError("rwaven", e)
        self.K = K.value
        self.kwn = kwn[:self.K].copy()
        self.kwnwi = kwnwi[:self.K].copy()
        self.amin, self.amax = amin.value, amax.value
        return self.kwn, self.kwnwi

    def set_waven(self, kwn, kwnwi, swrtr=1):
        self.kwn = np.ascontiguousarray(kwn, np.float64)
        self.kwnwi = np.ascontiguousarray(kwnwi, np.float64)
        self.K = len(self.kwn)
        self.swrtr = swrtr

    # --- A2/A3 --------------------------------------------------------------
    def precal(self):
        K = self.K
        self.elbg = np.zeros(self.elanz * self.nimn * K)
        self.relbg = np.zeros(max(self.relanz, 1) * self.nimn)
        self.kg = np.zeros(max(self.relanz, 1) * self.mesh.eanz * K)
        lb, lr, st = C.c_int(), C.c_int(), C.c_double()
        e = lib().orc_precal(C.byref(self.cm), K, self.kwn, self.elbg, self.relbg, self.kg,
                             C.byref(lb), C.byref(lr), C.byref(st))
        if e:
            raise OracleError("precal", e)
        self.lbeta, self.lrandb, self.sytop = lb.value, lr.value, st.value
        return self.elbg, self.relbg, self.kg

    def assemble(self, sigma, l, k):
        n, mb = self.mesh.sanz, self.mesh.mb
        a = np.zeros((mb + 1) * n, np.complex128)
        b = np.zeros(n, np.complex128)
        fak = np.zeros(n)
        e = lib().orc_assemble(C.byref(self.cm), self.K, self.elbg, self.relbg, self.kg,
                               self.lrandb, np.ascontiguousarray(sigma, np.complex128), l, k, a,
                               b, fak)
        if e:
            raise OracleError("assemble", e)
        return a, b, fak

    def chol(self, a):
        a = a.copy()
        e = lib().orc_chol(self.mesh.sanz, self.mesh.mb, a)
        if e:
            raise OracleError("chol", e)
        return a

    def vre(self, a, b):
        pot = np.zeros(self.mesh.sanz, np.complex128)
        lib().orc_vre(self.mesh.sanz, self.mesh.mb, a, np.ascontiguousarray(b), pot)
        return pot

    def solve_ld(self, sigma, l, k):
        """kpot(:, l, k) of the assembled system solved in long double (a
        verification aid: the exact solution of the discrete system to ~1e-19
        times its condition number)."""
        a, b, fak = self.assemble(sigma, l, k)
        pot = np.zeros(self.mesh.sanz, np.complex128)
        e = lib().orc_solve_ld(self.mesh.sanz, self.mesh.mb, a, b, pot)
        if e:
            raise OracleError("solve_ld", e)
        return pot * fak

    # --- A4..A11 ------------------------------------------------------------
    def forward(self, sigma, nthreads=0, sys_range=None):
        n, E, K = self.mesh.sanz, self.mesh.eanz, self.K
        kpot = np.zeros(n * E * K, np.complex128)
        hpot = np.zeros(n * E, np.complex128)
        b, e_ = (0, 0) if sys_range is None else sys_range
        e = lib().orc_forward(C.byref(self.cm), K, self.kwnwi, self.swrtr, self.elbg, self.relbg,
                              self.kg, self.lbeta, self.lrandb,
                              np.ascontiguousarray(sigma, np.complex128), kpot, hpot, nthreads,
                              b, e_)
        if e:
            raise OracleError("forward", e)
        self.kpot = kpot.reshape((n, E, K), order="F")
        self.hpot = hpot.reshape((n, E), order="F")
        self.sigma = np.ascontiguousarray(sigma, np.complex128)
        return self.kpot, self.hpot

    def forward_systems(self, sigma, systems, workers=None):
        """kpot columns of the listed (l, k) systems only (one factorisation
        each, as forward computes them), in parallel threads; the rest of
        self.kpot stays zero.  Returns self.kpot (sanz, eanz, K)."""
        import os
        from concurrent.futures import ThreadPoolExecutor
        n, E, K = self.mesh.sanz, self.mesh.eanz, self.K
        kpot = np.zeros(n * E * K, np.complex128)
        hpot = np.zeros(n * E, np.complex128)
        sig = np.ascontiguousarray(sigma, np.complex128)
        workers = workers or min(16, os.cpu_count() or 1)

        def one(lk):
            s = lk[1] * E + lk[0]
            return lib().orc_forward(C.byref(self.cm), K, self.kwnwi, self.swrtr, self.elbg, self.relbg,
                                     self.kg, self.lbeta, self.lrandb, sig, kpot, hpot, 1, s, s + 1)

        with ThreadPoolExecutor(max_workers=workers) as ex:
            errs = list(ex.map(one, list(systems)))
        for e in errs:
            if e:
                raise OracleError("forward", e)
        self.kpot = kpot.reshape((n, E, K), order="F")
        self.sigma = sig
        return self.kpot

    # --- A19: DC twins --------------------------------------------------------
    def assemble_dc(self, sigma, l, k):
        """kompadc + randdc + scaldc: the scaled real band matrix, rhs and fak."""
        n, mb = self.mesh.sanz, self.mesh.mb
        a = np.zeros((mb + 1) * n)
        b = np.zeros(n)
        fak = np.zeros(n)
        e = lib().orc_assemble_dc(C.byref(self.cm), self.elbg, self.relbg, self.kg, self.lrandb,
                                  np.ascontiguousarray(sigma, np.complex128), l, k, a, b, fak)
        if e:
            raise OracleError("assemble_dc", e)
        return a, b, fak

    def choldc(self, a):
        a = a.copy()
        e = lib().orc_choldc(self.mesh.sanz, self.mesh.mb, a)
        if e:
            raise OracleError("choldc", e)
        return a

    def vredc(self, a, b):
        pot = np.zeros(self.mesh.sanz)
        lib().orc_vredc(self.mesh.sanz, self.mesh.mb, a, np.ascontiguousarray(b, np.float64), pot)
        return pot

    def forward_dc(self, sigma, nthreads=0):
        """kpotdc (sanz, eanz, K) and hpotdc (sanz, eanz), real; the handle's
        hpot/kpot become their complex copies (dcmplx), as bvolti/bsendc read
        them in the DC case."""
        n, E, K = self.mesh.sanz, self.mesh.eanz, self.K
        kp = np.zeros(n * E * K)
        hp = np.zeros(n * E)
        e = lib().orc_forward_dc(C.byref(self.cm), K, self.kwnwi, self.swrtr, self.elbg, self.relbg,
                                 self.kg, self.lbeta, self.lrandb,
                                 np.ascontiguousarray(sigma, np.complex128), kp, hp, nthreads)
        if e:
            raise OracleError("forward_dc", e)
        self.kpotdc = kp.reshape((n, E, K), order="F")
        self.hpotdc = hp.reshape((n, E), order="F")
        self.kpot = self.kpotdc.astype(np.complex128)
        self.hpot = self.hpotdc.astype(np.complex128)
        self.sigma = np.ascontiguousarray(sigma, np.complex128)
        return self.kpotdc, self.hpotdc

    def forward_dc_systems(self, sigma, systems, workers=None):
        """kpotdc columns (sanz,) of the listed (l, k) systems, as
        orc_forward_dc computes them (crt_oracle.c, the DC loop of
        src/inv.f90:343-400): kompadc + randdc + scaldc, choldc, vredc,
        pot * fak; without mixed boundaries the sources l > 0 take kompbdc's
        vector with the factor of source 0.  Returns {(l, k): column}."""
        import os
        from concurrent.futures import ThreadPoolExecutor
        m = self.mesh
        sig = np.ascontiguousarray(sigma, np.complex128)
        workers = workers or min(16, os.cpu_count() or 1)

        def one(lk):
            l, k = lk
            a, b, fak = self.assemble_dc(sig, l if self.lbeta else 0, k)
            a = self.choldc(a)
            if not self.lbeta and l > 0:
                enr = self._keep["enr"]
                b = np.zeros(m.sanz)
                b[enr[l] - 1] = -fak[enr[l] - 1]
                if self.cm.lsink:
                    b[self.cm.nsink - 1] = fak[self.cm.nsink - 1]
            return self.vredc(a, b) * fak

        systems = list(systems)
        with ThreadPoolExecutor(max_workers=workers) as ex:
            cols = list(ex.map(one, systems))
        return dict(zip(systems, cols))

    def sens_dc(self, nthreads=0):
        """bsendc: real sensdc (nanz, elanz)."""
        N = len(self.strnr)
        out = np.zeros(N * self.elanz)
        kp = np.asfortranarray(self.kpotdc).ravel(order="F")
        e = lib().orc_sens_dc(C.byref(self.cm), N, self.strnr, self.vnr, self.K, self.kwnwi,
                              self.swrtr, self.elbg, kp, self.sigma, self.voltv, out, nthreads)
        if e:
            raise OracleError("sens_dc", e)
        return out.reshape((N, self.elanz), order="F")

    # --- A12 ----------------------------------------------------------------
    def volt(self, strnr, vnr, mode=0):
        N = len(strnr)
        volt = np.zeros(N, np.complex128)
        sigmaa = np.zeros(N, np.complex128)
        hp = np.asfortranarray(self.hpot).ravel(order="F")
        e = lib().orc_volt(C.byref(self.cm), N, np.ascontiguousarray(strnr, np.int32),
                           np.ascontiguousarray(vnr, np.int32), hp, mode, None, volt, sigmaa)
        if e:
            raise OracleError("volt", e)
        self.strnr = np.ascontiguousarray(strnr, np.int32)
        self.vnr = np.ascontiguousarray(vnr, np.int32)
        self.voltv = volt
        return volt, sigmaa

    # --- A14 ----------------------------------------------------------------
    def sens(self, mode=0, nthreads=0):
        N = len(self.strnr)
        out = np.zeros(N * self.elanz, np.complex128)
        kp = np.asfortranarray(self.kpot).ravel(order="F")
        e = lib().orc_sens(C.byref(self.cm), N, self.strnr, self.vnr, self.K, self.kwnwi,
                           self.swrtr, self.elbg, kp, mode, self.sigma, self.voltv, out, nthreads)
        if e:
            raise OracleError("sens", e)
        return out.reshape((N, self.elanz), order="F")


def bsmatmreg(orc: Oracle, nx, nz, alfx, alfz):
    sm = np.zeros(orc.elanz * 3)
    lib().orc_bsmatmreg(C.byref(orc.cm), nx, nz, alfx, alfz, sm)
    return sm.reshape((orc.elanz, 3), order="F")


def bnachbar(orc: Oracle):
    """Edge neighbours (manz, maxnodes+1), 1-based, last column = count."""
    mx = orc.mesh.maxnodes
    nb = np.zeros(orc.elanz * (mx + 1), np.int32)
    lib().orc_bnachbar(C.byref(orc.cm), nb)
    return nb.reshape((orc.elanz, mx + 1), order="F")


def bsmatmtri(orc: Oracle, alfx, alfz, nachbar, decouplings=None):
    """Triangular smoothness operator smatm (manz, maxnodes+1).
    decouplings: (edecoup (ndec, 2) 1-based element pairs, edecstr (ndec,))."""
    mx = orc.mesh.maxnodes
    sm = np.zeros(orc.elanz * (mx + 1))
    if decouplings is None or len(decouplings[1]) == 0:
        ed, es = np.zeros(2, np.int32), np.zeros(1)
        nd = 0
    else:
        ed = np.asfortranarray(decouplings[0], np.int32).ravel(order="F").copy()
        es = np.ascontiguousarray(decouplings[1], np.float64)
        nd = len(es)
    lib().orc_bsmatmtri(C.byref(orc.cm), alfx, alfz,
                        np.asfortranarray(nachbar, np.int32).ravel(order="F"), nd, ed, es, sm)
    return sm.reshape((orc.elanz, mx + 1), order="F")


def rm_reg(smatm, nx, par):
    manz = smatm.shape[0]
    out = np.zeros(manz, np.complex128)
    lib().orc_rm_reg(manz, nx, np.asfortranarray(smatm).ravel(order="F"),
                     np.ascontiguousarray(par, np.complex128), out)
    return out


def update_rhs(sens, wmatd, wdfak, dat, sigmaa, lam, ltri, rdiag, rm=None):
    N, M = sens.shape
    cgfac = np.zeros(M)
    bvec = np.zeros(M, np.complex128)
    rm_arr = None if rm is None else np.ascontiguousarray(rm, np.complex128)
    lib().orc_update_rhs(N, M, np.asfortranarray(sens).ravel(order="F"),
                         np.ascontiguousarray(wmatd, np.float64),
                         np.ascontiguousarray(wdfak, np.int32),
                         np.ascontiguousarray(dat, np.complex128),
                         np.ascontiguousarray(sigmaa, np.complex128), lam, ltri,
                         np.ascontiguousarray(rdiag, np.float64),
                         None if rm_arr is None else rm_arr.ctypes.data, cgfac, bvec)
    return cgfac, bvec


def cjggra(sens, wmatd, wdfak, cgfac, bvec, lam, ltri, nx, smatm, ncgmax, eps, nachbar=None,
           maxnodes=0):
    N, M = sens.shape
    dpar = np.zeros(M, np.complex128)
    cgres = np.zeros(ncgmax + 2)
    sm = np.asfortranarray(smatm)
    nb = None if nachbar is None else np.asfortranarray(nachbar, np.int32)
    ncg = lib().orc_cjggra(N, M, np.asfortranarray(sens).ravel(order="F"),
                           np.ascontiguousarray(wmatd, np.float64),
                           np.ascontiguousarray(wdfak, np.int32),
                           np.ascontiguousarray(cgfac, np.float64),
                           np.ascontiguousarray(bvec, np.complex128), lam, ltri, nx,
                           sm.ravel(order="F"), sm.shape[0], None if nb is None else nb.ctypes.data,
                           maxnodes, ncgmax, eps, dpar, cgres)
    return dpar, cgres, ncg


def adjoint_grad(sens, wmatd, wdfak, resid):
    N, M = sens.shape
    g = np.zeros(M, np.complex128)
    lib().orc_adjoint_grad(N, M, np.asfortranarray(sens).ravel(order="F"),
                           np.ascontiguousarray(wmatd, np.float64),
                           np.ascontiguousarray(wdfak, np.int32),
                           np.ascontiguousarray(resid, np.complex128), g)
    return g


def nthreads_default():
    return int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))


# --- inversion-loop pieces: dmisft / chkpo2, bpar, bsigma, update step --------

def _inv_lib():
    L = lib()
    if not getattr(L, "_inv_bound", False):
        L.orc_chkpo2.argtypes = [_OrcC, _OrcC, C.c_int, C.POINTER(_OrcC), C.POINTER(_OrcC),
                                 C.POINTER(C.c_int)]
        L.orc_dmisft.argtypes = [C.c_int, _z, _z, _d, _d, _d, C.c_void_p, C.c_void_p, C.c_int,
                                 C.c_double, _d, _i]
        L.orc_bpar.argtypes = [C.c_int, C.c_void_p, _z, _z]
        L.orc_bsigma.argtypes = [C.c_int, C.c_void_p, _z, _z]
        L.orc_update_step.argtypes = [C.c_int, _z, _z, C.c_double, C.c_int, _d]
        L._inv_bound = True
    return L


class _OrcC(C.Structure):
    _fields_ = [("re", C.c_double), ("im", C.c_double)]


def chkpo2(dati, sigi, ldum):
    """(cdat, csig, wdlok) of src/dmisft.f90:213-300 for one datum."""
    cd, cs, wd = _OrcC(), _OrcC(), C.c_int(1)
    _inv_lib().orc_chkpo2(_OrcC(dati.real, dati.imag), _OrcC(sigi.real, sigi.imag), int(bool(ldum)),
                          C.byref(cd), C.byref(cs), C.byref(wd))
    return complex(cd.re, cd.im), complex(cs.re, cs.im), wd.value


DMISFT_LPOL, DMISFT_LFPI, DMISFT_LDIFF, DMISFT_LROBUST = 1, 2, 4, 8


def dmisft(dat, sigmaa, wmatd, wmatdr, wmatdp, flags=0, l1min=1.0, d0=None, fm0=None):
    """Returns (stats, wmatd_out, wdlok); stats = nrmsd, betrms, pharms,
    rmssum, l1rat, npol."""
    N = len(dat)
    w = np.array(wmatd, np.float64)
    st = np.zeros(6)
    wd = np.zeros(N, np.int32)
    d0p = np.ascontiguousarray(d0, np.complex128) if d0 is not None else None
    fmp = np.ascontiguousarray(fm0, np.complex128) if fm0 is not None else None
    e = _inv_lib().orc_dmisft(N, np.ascontiguousarray(dat, np.complex128),
                              np.ascontiguousarray(sigmaa, np.complex128), w,
                              np.ascontiguousarray(wmatdr, np.float64), np.ascontiguousarray(wmatdp, np.float64),
                              d0p.ctypes.data if d0p is not None else None,
                              fmp.ctypes.data if fmp is not None else None, int(flags), float(l1min), st, wd)
    if e:
        raise OracleError("dmisft", e)
    return st, w, wd


def bpar(sigma, manz=None, mnr=None, par=None):
    sigma = np.ascontiguousarray(sigma, np.complex128)
    manz = len(sigma) if manz is None else manz
    out = np.zeros(manz, np.complex128) if par is None else np.array(par, np.complex128)
    m = np.ascontiguousarray(mnr, np.int32) if mnr is not None else None
    _inv_lib().orc_bpar(len(sigma), m.ctypes.data if m is not None else None, sigma, out)
    return out


def bsigma(par, elanz=None, mnr=None):
    par = np.ascontiguousarray(par, np.complex128)
    elanz = len(par) if elanz is None else elanz
    out = np.zeros(elanz, np.complex128)
    m = np.ascontiguousarray(mnr, np.int32) if mnr is not None else None
    _inv_lib().orc_bsigma(elanz, m.ctypes.data if m is not None else None, par, out)
    return out


def update_step(par, dpar, step, lphi0=False):
    """(par_new, bdpar, forced) of src/update.f90:332-351."""
    p = np.array(par, np.complex128)
    out = np.zeros(2)
    _inv_lib().orc_update_step(len(p), p, np.ascontiguousarray(dpar, np.complex128), float(step),
                               int(bool(lphi0)), out)
    return p, out[0], int(out[1])


def blam0(sens, w, mode=0, alfx=1.0, alfz=1.0):
    """Starting lambda of src/blam0.f90 (reference loop order, O(N M^2));
    mode 0 complex, 1 real parts (FPI / DC)."""
    L = lib()
    L.orc_blam0.restype = C.c_double
    L.orc_blam0.argtypes = [C.c_int, C.c_int, _z, _d, C.c_int, C.c_double, C.c_double]
    S = np.asfortranarray(sens, np.complex128)
    N, M = S.shape
    return L.orc_blam0(N, M, np.ascontiguousarray(S.ravel(order="F")), np.ascontiguousarray(w, np.float64),
                       int(mode), float(alfx), float(alfz))
