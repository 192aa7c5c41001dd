"""ctypes binding of the reference harness (oracle/_ref/libcrtref.so).

TEST INFRASTRUCTURE ONLY.  The library holds the reference's own Fortran routines
compiled from its source by oracle/Makefile plus oracle/ref_driver.f90.  It is
module-global state (like the reference program), so one mesh at a time.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
LIB = HERE / "_ref" / "libcrtref.so"

_i = np.ctypeslib.ndpointer(np.int32, flags="C")
_d = np.ctypeslib.ndpointer(np.float64, flags="C")
_z = np.ctypeslib.ndpointer(np.complex128, flags="C")

_lib = None


def available() -> bool:
    return LIB.exists()


def lib():
    global _lib
    if _lib is None:
        L = C.CDLL(str(LIB))
        L.ref_load_mesh.argtypes = [C.c_int, C.c_int, C.c_int, _i, _i, _i, _i, _i, _i, _d, _d,
                                    C.c_int, _i, C.c_int, C.c_int, C.c_int]
        L.ref_load_data.argtypes = [C.c_int, _i, _i]
        L.ref_rwaven.argtypes = [C.POINTER(C.c_int), _d, _d, C.POINTER(C.c_double),
                                 C.POINTER(C.c_double)]
        L.ref_set_waven.argtypes = [C.c_int, _d, _d, C.c_int]
        L.ref_precal.argtypes = [C.POINTER(C.c_int), C.POINTER(C.c_int), _d, _d, _d,
                                 C.POINTER(C.c_double)]
        L.ref_assemble.argtypes = [C.c_int, C.c_int, _z, _z, _z, _d]
        L.ref_chol.argtypes = [_z]
        L.ref_vre.argtypes = [_z, _z, _z]
        L.ref_forward.argtypes = [_z, _z, _z, C.c_int]
        L.ref_volt.argtypes = [C.c_int, _z, _z]
        L.ref_sens.argtypes = [C.c_int, _z]
        L.ref_forward_dc.argtypes = [_z, _d, _d, C.c_int]
        L.ref_set_randb2.argtypes = [C.c_int, _i, _z]
        L.ref_bsmatmtri.argtypes = [C.c_double, C.c_double, C.c_int, _i, _d, _d, _i]
        L.ref_bsmatmreg.argtypes = [C.c_int, C.c_int, C.c_double, C.c_double, _d]
        L.ref_assemble_dc.argtypes = [C.c_int, C.c_int, _z, _d, _d, _d]
        L.ref_choldc.argtypes = [_d]
        L.ref_sens_dc.argtypes = [_d]
        _lib = L
    return _lib


class RefError(RuntimeError):
    def __init__(self, where, code):
        super().__init__(f"reference {where}: errnr {code}")
        self.code = code


class Reference:
    """Drive the reference routines for one mesh (module-global state)."""

    def __init__(self, mesh, lsytop=False, lsink=False, nsink=0):
        self.mesh = mesh
        nrel = np.asfortranarray(mesh.nrel, np.int32).ravel(order="F").copy()
        rnr = np.ascontiguousarray(mesh.rnr if len(mesh.rnr) else np.zeros(1), np.int32)
        e = lib().ref_load_mesh(mesh.sanz, mesh.typanz, mesh.mb,
                                np.ascontiguousarray(mesh.typ, np.int32),
                                np.ascontiguousarray(mesh.nelanz, np.int32),
                                np.ascontiguousarray(mesh.selanz, np.int32), nrel, rnr,
                                np.ascontiguousarray(mesh.snr, np.int32),
                                np.ascontiguousarray(mesh.sx, np.float64),
                                np.ascontiguousarray(mesh.sy, np.float64), mesh.eanz,
                                np.ascontiguousarray(mesh.enr, np.int32), int(lsytop),
                                int(lsink), int(nsink))
        if e:
            raise RefError("load_mesh", e)
        self.elanz, self.relanz = mesh.elanz, mesh.relanz
        self.nimn = mesh.maxnodes * (mesh.maxnodes + 1) // 2

    def load_data(self, strnr, vnr):
        self.nanz = len(strnr)
        lib().ref_load_data(self.nanz, np.ascontiguousarray(strnr, np.int32),
                            np.ascontiguousarray(vnr, np.int32))

    def set_boundary_values(self, rwdnr, rwd):
        rn = np.ascontiguousarray(rwdnr, np.int32)
        rv = np.ascontiguousarray(rwd, np.complex128)
        lib().ref_set_randb2(len(rn), rn if len(rn) else np.zeros(1, np.int32),
                             rv if len(rv) else np.zeros(1, np.complex128))

    def rwaven(self):
        K = C.c_int()
        kwn = np.zeros(30)
        kwnwi = np.zeros(30)
        amin, amax = C.c_double(), C.c_double()
        e = lib().ref_rwaven(C.byref(K), kwn, kwnwi, C.byref(amin), C.byref(amax))
        if e:
            raise RefError("rwaven", e)
        self.K = K.value
        self.kwn, self.kwnwi = kwn[:self.K].copy(), kwnwi[:self.K].copy()
        return self.kwn, self.kwnwi

    def set_waven(self, kwn, kwnwi, swrtr=1):
        self.K = len(kwn)
        lib().ref_set_waven(self.K, np.ascontiguousarray(kwn, np.float64),
                            np.ascontiguousarray(kwnwi, np.float64), swrtr)

    def precal(self):
        K = self.K
        elbg = np.zeros(self.elanz * self.nimn * K)
        relbg = np.zeros(max(self.relanz, 1) * self.nimn)
        kg = np.zeros(max(self.relanz, 1) * self.mesh.eanz * K)
        lb, lr, st = C.c_int(), C.c_int(), C.c_double()
        e = lib().ref_precal(C.byref(lb), C.byref(lr), elbg, relbg, kg, C.byref(st))
        if e:
            raise RefError("precal", e)
        self.lbeta, self.lrandb, self.sytop = lb.value, lr.value, st.value
        return elbg, relbg, kg

    def assemble(self, sigma, l, k):
        n, mb = self.mesh.sanz, self.mesh.mb
        a = np.zeros((mb + 1) * n, np.complex128)
        b = np.zeros(n, np.complex128)
        fak = np.zeros(n)
        e = lib().ref_assemble(l + 1, k + 1, np.ascontiguousarray(sigma, np.complex128), a, b, fak)
        if e:
            raise RefError("assemble", e)
        return a, b, fak

    def chol(self, a):
        a = a.copy()
        e = lib().ref_chol(a)
        if e:
            raise RefError("chol", e)
        return a

    def vre(self, a, b):
        pot = np.zeros(self.mesh.sanz, np.complex128)
        lib().ref_vre(a, np.ascontiguousarray(b), pot)
        return pot

    def forward(self, sigma, nthreads=0):
        n, E, K = self.mesh.sanz, self.mesh.eanz, self.K
        kpot = np.zeros(n * E * K, np.complex128)
        hpot = np.zeros(n * E, np.complex128)
        e = lib().ref_forward(np.ascontiguousarray(sigma, np.complex128), kpot, hpot, nthreads)
        if e:
            raise RefError("forward", e)
        return kpot.reshape((n, E, K), order="F"), hpot.reshape((n, E), order="F")

    def volt(self, mode=0):
        volt = np.zeros(self.nanz, np.complex128)
        sigmaa = np.zeros(self.nanz, np.complex128)
        e = lib().ref_volt(mode, volt, sigmaa)
        if e:
            raise RefError("volt", e)
        return volt, sigmaa

    def sens(self, mode=0):
        out = np.zeros(self.nanz * self.elanz, np.complex128)
        e = lib().ref_sens(mode, out)
        if e:
            raise RefError("sens", e)
        return out.reshape((self.nanz, self.elanz), order="F")

    # DC twins (ldc): kompadc/randdc/scaldc/choldc/vredc/kompbdc, rtrafo's DC
    # branch, bvolti's DC branch (through volt(1)) and bsendc
    def forward_dc(self, sigma, nthreads=0):
        n, E, K = self.mesh.sanz, self.mesh.eanz, self.K
        kp = np.zeros(n * E * K)
        hp = np.zeros(n * E)
        e = lib().ref_forward_dc(np.ascontiguousarray(sigma, np.complex128), kp, hp, nthreads)
        if e:
            raise RefError("forward_dc", e)
        return kp.reshape((n, E, K), order="F"), hp.reshape((n, E), order="F")

    def assemble_dc(self, sigma, l, k):
        n, mb = self.mesh.sanz, self.mesh.mb
        a = np.zeros((mb + 1) * n)
        b = np.zeros(n)
        fak = np.zeros(n)
        e = lib().ref_assemble_dc(l + 1, k + 1, np.ascontiguousarray(sigma, np.complex128), a, b, fak)
        if e:
            raise RefError("assemble_dc", e)
        return a, b, fak

    def choldc(self, a):
        a = a.copy()
        e = lib().ref_choldc(a)
        if e:
            raise RefError("choldc", e)
        return a

    def bsmatmtri(self, alfx, alfz, decouplings=None):
        """bnachbar + bsmatm(ltri=1): smatm and nachbar, (manz, maxnodes+1).
        decouplings: (edecoup (ndec, 2) 1-based element pairs, edecstr (ndec,))."""
        mx = self.mesh.maxnodes
        sm = np.zeros(self.elanz * (mx + 1))
        nb = np.zeros(self.elanz * (mx + 1), np.int32)
        if decouplings is None or len(decouplings[1]) == 0:
            ed, es, nd = np.zeros(2, np.int32), np.zeros(1), 0
        else:
            ed = np.asfortranarray(decouplings[0], np.int32).ravel(order="F").copy()
            es = np.ascontiguousarray(decouplings[1], np.float64)
            nd = len(es)
        e = lib().ref_bsmatmtri(alfx, alfz, nd, ed, es, sm, nb)
        if e:
            raise RefError("bsmatmtri", e)
        return sm.reshape((self.elanz, mx + 1), order="F"), nb.reshape((self.elanz, mx + 1), order="F")

    def bsmatmreg(self, nx, nz, alfx, alfz):
        """bsmatm(ltri=0), i.e. bsmatmreg: smatm (manz, 3)."""
        sm = np.zeros(self.elanz * 3)
        e = lib().ref_bsmatmreg(int(nx), int(nz), float(alfx), float(alfz), sm)
        if e:
            raise RefError("bsmatmreg", e)
        return sm.reshape((self.elanz, 3), order="F")

    def sens_dc(self):
        out = np.zeros(self.nanz * self.elanz)
        e = lib().ref_sens_dc(out)
        if e:
            raise RefError("sens_dc", e)
        return out.reshape((self.nanz, self.elanz), order="F")


def forward_sample(sigma, first, last, nthreads=0):
    """Time-only sample of the reference forward loop (systems first..last, 1-based)."""
    L = lib()
    L.ref_forward_sample.argtypes = [_z, C.c_int, C.c_int, C.c_int]
    e = L.ref_forward_sample(np.ascontiguousarray(sigma, np.complex128), first, last, nthreads)
    if e:
        raise RefError("forward_sample", e)


# --- inversion-loop pieces (src/dmisft.f90, src/bpar.f90, src/bsigma.f90) ----

def chkpo2(dati, sigi, ldum):
    L = lib()
    L.ref_chkpo2.argtypes = [_z, _z, C.c_int, _z, _z, _i]
    cd = np.zeros(1, np.complex128)
    cs = np.zeros(1, np.complex128)
    wd = np.zeros(1, np.int32)
    L.ref_chkpo2(np.array([dati], np.complex128), np.array([sigi], np.complex128), int(bool(ldum)), cd, cs, wd)
    return complex(cd[0]), complex(cs[0]), int(wd[0])


def dmisft(dat, sigmaa, wmatd, wmatdr, wmatdp, flags=0, l1min=1.0, d0=None, fm0=None, workdir=None):
    """The reference dmisft; its control files go to ``workdir`` (a short
    temporary directory by default).  Returns (stats, wmatd_out) with stats =
    nrmsd, betrms, pharms, rmssum, l1rat, npol."""
    import tempfile
    L = lib()
    L.ref_dmisft.argtypes = [C.c_int, _z, _z, _d, _d, _d, _z, _z, C.c_int, C.c_double, C.c_char_p,
                             C.c_int, _d]
    N = len(dat)
    z = np.zeros(N, np.complex128)
    w = np.array(wmatd, np.float64)
    out = np.zeros(6)
    with tempfile.TemporaryDirectory(dir="/tmp", prefix="crtr") as td:
        d = (workdir or td).encode()
        assert len(d) < 60
        e = L.ref_dmisft(N, np.ascontiguousarray(dat, np.complex128), np.ascontiguousarray(sigmaa, np.complex128),
                         w, np.ascontiguousarray(wmatdr, np.float64), np.ascontiguousarray(wmatdp, np.float64),
                         np.ascontiguousarray(d0 if d0 is not None else z, np.complex128),
                         np.ascontiguousarray(fm0 if fm0 is not None else z, np.complex128),
                         int(flags), float(l1min), d, len(d), out)
    if e:
        raise RefError("dmisft", e)
    return out, w


def bpar(ref_mesh, sigma, manz, mnr, par=None):
    """bpar for the loaded mesh (a Reference instance holds it)."""
    del ref_mesh
    L = lib()
    L.ref_bpar.argtypes = [C.c_int, _i, _z, _z]
    out = np.zeros(manz, np.complex128) if par is None else np.array(par, np.complex128)
    e = L.ref_bpar(manz, np.ascontiguousarray(mnr, np.int32), np.ascontiguousarray(sigma, np.complex128), out)
    if e:
        raise RefError("bpar", e)
    return out


def bsigma(ref_mesh, par, elanz, mnr):
    del ref_mesh
    L = lib()
    L.ref_bsigma.argtypes = [C.c_int, _i, _z, _z]
    out = np.zeros(elanz, np.complex128)
    e = L.ref_bsigma(len(par), np.ascontiguousarray(mnr, np.int32), np.ascontiguousarray(par, np.complex128), out)
    if e:
        raise RefError("bsigma", e)
    return out


def blam0(sens, w, mode=0, alfx=1.0, alfz=1.0):
    """The reference blam0 on J (N x M): mode 0 complex, 1 lfpi, 2 ldc."""
    L = lib()
    L.ref_blam0.argtypes = [C.c_int, C.c_int, _z, _d, C.c_int, C.c_double, C.c_double, C.POINTER(C.c_double)]
    S = np.asfortranarray(sens, np.complex128)
    N, M = S.shape
    out = C.c_double()
    L.ref_blam0(N, M, np.ascontiguousarray(S.ravel(order="F")), np.ascontiguousarray(w, np.float64), int(mode),
                float(alfx), float(alfz), C.byref(out))
    return out.value


def rdati(path, eanz, ldc=False, model=(5.0, 1e-3, 0.0, 0.0, 0.0, 0.0), nmax=100000, noise=None):
    """The reference rdati on a data file: (errnr, dict of arrays).  noise =
    (iseed, nstabw0, nstabm0, nstabpA1, nstabpB, nstabpA2, nstabp0) switches
    on its synthetic noise; its listing files go to a temporary directory."""
    import os
    import tempfile
    L = lib()
    L.ref_rdati.argtypes = [C.c_char_p, C.c_int, C.c_int, C.c_int, _d, C.c_int, _i, _i, _i, _z, _d, _d, _d, _i,
                            C.c_int, _d]
    p = str(Path(path).resolve()).encode()
    nz = np.asarray(noise if noise is not None else (1, 0, 0, 0, 0, 0, 0), np.float64)
    info = np.zeros(2, np.int32)
    st = np.zeros(nmax, np.int32)
    vn = np.zeros(nmax, np.int32)
    wd = np.zeros(nmax, np.int32)
    dat = np.zeros(nmax, np.complex128)
    wr = np.zeros(nmax)
    wp = np.zeros(nmax)
    wc = np.zeros(nmax)
    cwd = os.getcwd()
    with tempfile.TemporaryDirectory(dir="/tmp", prefix="crtn") as td:
        os.chdir(td)
        try:
            e = L.ref_rdati(p, len(p), int(eanz), int(bool(ldc)), np.asarray(model, np.float64), nmax, info, st, vn,
                            dat, wr, wp, wc, wd, int(noise is not None), nz)
        finally:
            os.chdir(cwd)
    n = int(info[0])
    return e, dict(strnr=st[:n], vnr=vn[:n], dat=dat[:n], wmatdr=wr[:n], wmatdp=wp[:n], wmatd_cri=wc[:n],
                   wdfak=wd[:n], lindiv=bool(info[1]))


def update(sens, w, dat, sigmaa, par, smatm, lam, nx, nz, ncgmax, eps, step=1.0):
    """The reference update (llam = .false., ltri = 0): (dpar, cgfac, cgres,
    par_new, bdpar); dpar is scaled by cgfac as update leaves it."""
    L = lib()
    L.ref_update.argtypes = [C.c_int, C.c_int, _z, _d, _z, _z, _z, _d, C.c_double, C.c_int, C.c_int, C.c_int,
                             C.c_double, C.c_double, _z, _d, _d, C.POINTER(C.c_double)]
    S = np.asfortranarray(sens, np.complex128)
    N, M = S.shape
    p = np.array(par, np.complex128)
    dpar = np.zeros(M, np.complex128)
    cgfac = np.zeros(M)
    cgres = np.zeros(ncgmax + 1)
    bd = C.c_double()
    e = L.ref_update(N, M, np.ascontiguousarray(S.ravel(order="F")), np.ascontiguousarray(w, np.float64),
                     np.ascontiguousarray(dat, np.complex128), np.ascontiguousarray(sigmaa, np.complex128), p,
                     np.ascontiguousarray(np.asfortranarray(smatm, np.float64).ravel(order="F")), float(lam),
                     int(nx), int(nz), int(ncgmax), float(eps), float(step), dpar, cgfac, cgres, C.byref(bd))
    if e:
        raise RefError("update", e)
    return dpar, cgfac, cgres, p, bd.value
