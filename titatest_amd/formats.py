"""CRMod/CRTomo on-disk formats: elem.dat, elec.dat, rho.dat, config.dat, volt.dat.

Readers and writers with the semantics of the reference's own I/O routines, so a
tomodir prepared for CRMod reads the same here:

* ``elem.dat``  -- src/relem.f90:62-151 (layout), :163-252 (border checks and the
  ``rnr`` repair that CRMod applies in memory).
* ``elec.dat``  -- src/relectr.f90:46-68.
* ``rho.dat``   -- src/rsigma.f90:91-189; sigma = (cos(phi)/|rho|, -sin(phi)/|rho|)
  with phi given in mrad (src/rsigma.f90:187).
* ``config.dat`` -- src/rdatm.f90:54-115 (CRTomo ``A*10^4+B`` form or ``A B M N``).
* ``volt.dat``  -- src/wdatm.f90:55-79 (writer, including the polarity swap of the
  potential pair when the phase is near +-pi); ``read_volt`` gives the columns
  and ``read_data`` reads a measured data file the way CRTomo's rdati does
  (src/rdati.f90:84-363: format detection, individual errors or the error
  model, validation with the reference's error numbers, the data vector and
  weights).

All node / element / electrode numbers are kept 1-based, as in the files and in
the Fortran module arrays the C-ABI mirrors.
"""
from __future__ import annotations

import dataclasses
import math
from pathlib import Path

import numpy as np


@dataclasses.dataclass
class Mesh:
    """Finite-element mesh in the reference's storage layout (src/elemmod.f90).

    ``nrel`` is (elanz + relanz) x maxnodes with 1-based node numbers (unused
    trailing slots are 0); element order is the file order: area element types
    first, boundary types (11 mixed, 12 no-flow, 13 Dirichlet) last.
    """

    sanz: int
    mb: int
    typ: np.ndarray          # (typanz,) int32
    nelanz: np.ndarray       # (typanz,) int32
    selanz: np.ndarray       # (typanz,) int32
    snr: np.ndarray          # (sanz,) int32, pointer to coordinates (1-based)
    sx: np.ndarray           # (sanz,) float64
    sy: np.ndarray           # (sanz,) float64
    nrel: np.ndarray         # (elanz+relanz, maxnodes) int32, 1-based
    rnr: np.ndarray          # (relanz,) int32, adjacent area element (1-based)
    enr: np.ndarray | None = None   # (eanz,) int32, electrode nodes (1-based)

    @property
    def typanz(self) -> int:
        return int(len(self.typ))

    @property
    def maxnodes(self) -> int:
        return int(self.selanz.max())

    @property
    def elanz(self) -> int:
        return int(self.nelanz[self.typ <= 10].sum())

    @property
    def relanz(self) -> int:
        return int(self.nelanz[self.typ > 10].sum())

    @property
    def eanz(self) -> int:
        return 0 if self.enr is None else int(len(self.enr))

    def element_ranges(self):
        """Yield (typ, nkel, first, last) 0-based element index ranges per type."""
        iel = 0
        for t, n, s in zip(self.typ, self.nelanz, self.selanz):
            yield int(t), int(s), iel, iel + int(n)
            iel += int(n)

    def node_xy(self, node1):
        """Coordinates of 1-based node(s), through the snr pointer."""
        idx = self.snr[np.asarray(node1) - 1] - 1
        return self.sx[idx], self.sy[idx]

    def check_bandwidth(self) -> int:
        """Largest |node difference| inside any element (kompab errnr 19 if > mb)."""
        worst = 0
        for t, nk, a, b in self.element_ranges():
            if t > 11:
                continue
            blk = self.nrel[a:b, :nk].astype(np.int64)
            worst = max(worst, int((blk.max(1) - blk.min(1)).max()) if b > a else 0)
        return worst


def _tokens(path):
    return Path(path).read_text().split()


def read_elem(path) -> Mesh:
    """Read elem.dat (src/relem.f90:62-151), then repair ``rnr`` the way relem does
    (src/relem.f90:209-238): each boundary element is linked to the area element
    holding both its nodes, searched over the first element type's node count."""
    tok = _tokens(path)
    p = 0

    def take(n, conv=int):
        nonlocal p
        out = [conv(v) for v in tok[p:p + n]]
        p += n
        return out

    sanz, typanz, mb = take(3)
    trip = np.array(take(3 * typanz), dtype=np.int64).reshape(typanz, 3)
    typ, nelanz, selanz = (trip[:, i].astype(np.int32) for i in range(3))
    nodes = take(3 * sanz, float)
    arr = np.array(nodes).reshape(sanz, 3)
    snr = arr[:, 0].astype(np.int32)
    sx = arr[:, 1].copy()
    sy = arr[:, 2].copy()
    maxn = int(selanz.max())
    nel = int(nelanz.sum())
    nrel = np.zeros((nel, maxn), dtype=np.int32)
    iel = 0
    for t, n, s in zip(typ, nelanz, selanz):
        blk = np.array(take(int(n) * int(s)), dtype=np.int32).reshape(int(n), int(s))
        nrel[iel:iel + n, :s] = blk
        iel += int(n)
    relanz = int(nelanz[typ > 10].sum())
    rnr = np.array(take(relanz), dtype=np.int32)
    mesh = Mesh(sanz=sanz, mb=mb, typ=typ, nelanz=nelanz, selanz=selanz, snr=snr,
                sx=sx, sy=sy, nrel=nrel, rnr=rnr)
    _repair_rnr(mesh)
    return mesh


def _repair_rnr(mesh: Mesh) -> int:
    """relem's 'check 2' (src/relem.f90:209-238), vectorised.  Returns #repairs."""
    elanz, relanz = mesh.elanz, mesh.relanz
    if relanz == 0:
        return 0
    s1 = int(mesh.selanz[0])
    area = mesh.nrel[:elanz, :s1]
    # map node -> elements containing it, using the first type's node count
    owners: dict[int, list[int]] = {}
    for k in range(elanz):
        for v in area[k]:
            owners.setdefault(int(v), []).append(k)
    fixed = 0
    for i in range(relanz):
        ik1, ik2 = (int(v) for v in mesh.nrel[elanz + i, :2])
        cand = set(owners.get(ik1, ())) & set(owners.get(ik2, ()))
        # relem keeps the LAST matching element in ascending k order
        for k in sorted(cand):
            if mesh.rnr[i] != k + 1:
                mesh.rnr[i] = k + 1
                fixed += 1
    bad = (mesh.rnr < 1) | (mesh.rnr > elanz)
    if bad.any():
        raise ValueError("border to element context is wrong (relem errnr 111)")
    return fixed


def write_elem(path, mesh: Mesh) -> None:
    lines = [f"{mesh.sanz:12d}{mesh.typanz:12d}{mesh.mb:12d}"]
    for t, n, s in zip(mesh.typ, mesh.nelanz, mesh.selanz):
        lines.append(f"{int(t):12d}{int(n):12d}{int(s):12d}")
    for i in range(mesh.sanz):
        lines.append(f"{int(mesh.snr[i]):8d} {mesh.sx[i]:24.16e} {mesh.sy[i]:24.16e}")
    for t, nk, a, b in mesh.element_ranges():
        for row in mesh.nrel[a:b, :nk]:
            lines.append(" ".join(f"{int(v):9d}" for v in row))
    for v in mesh.rnr:
        lines.append(f"{int(v):9d}")
    Path(path).write_text("\n".join(lines) + "\n")


def read_elec(path) -> np.ndarray:
    tok = _tokens(path)
    eanz = int(tok[0])
    enr = np.array([int(v) for v in tok[1:1 + eanz]], dtype=np.int32)
    return enr


def write_elec(path, enr) -> None:
    Path(path).write_text(f"{len(enr)}\n" + "".join(f"{int(v)}\n" for v in enr))


def read_rho(path, elanz: int | None = None) -> np.ndarray:
    """rho.dat -> complex conductivity per element (src/rsigma.f90:187).

    The optional reference-model flags after the element count are ignored the
    way rsigma does when they are absent (one integer on the first line)."""
    lines = [ln.split() for ln in Path(path).read_text().splitlines() if ln.strip()]
    n = int(lines[0][0])
    if elanz is not None and n != elanz:
        raise ValueError(f"rho.dat has {n} cells, mesh has {elanz} (rsigma errnr 47)")
    vals = np.array([[float(x) for x in ln[:2]] for ln in lines[1:1 + n]])
    bet = vals[:, 0]
    pha = 1e-3 * vals[:, 1]
    if (bet < 1e-12).any():
        raise ValueError("resistivity below 1e-12 (rsigma errnr 11)")
    return np.cos(pha) / bet - 1j * np.sin(pha) / bet


def write_rho(path, mag, pha_mrad) -> None:
    body = "".join(f"{m:.16e} {p:.16e}\n" for m, p in zip(mag, pha_mrad))
    Path(path).write_text(f"{len(mag)}\n" + body)


def read_config(path, eanz: int | None = None):
    """config.dat -> (strnr, vnr) int32 arrays (src/rdatm.f90:54-115)."""
    lines = [ln.split() for ln in Path(path).read_text().splitlines() if ln.strip()]
    nanz = int(lines[0][0])
    first = int(lines[1][0])
    crtf = first - 10000 >= 0
    strnr = np.empty(nanz, np.int32)
    vnr = np.empty(nanz, np.int32)
    for i in range(nanz):
        row = [int(float(v)) for v in lines[1 + i]]
        if crtf:
            strnr[i], vnr[i] = row[0], row[1]
        else:
            strnr[i] = row[0] * 10000 + row[1]
            vnr[i] = row[2] * 10000 + row[3]
    check_config(strnr, vnr, eanz)
    return strnr, vnr


def read_decouplings(path):
    """decouplings.dat -> (edecoup (decanz, 2) int32 1-based element pairs,
    edecstr (decanz,) float64 weight factors), as rall reads the file
    (src/rall.f90:648-665): the count, then one ``elem1 elem2 factor`` record
    per line.  The smoothness operator multiplies the weight of the edge
    between the two elements by the factor, in either orientation
    (src/bsmatm_mod.f90:434-445)."""
    lines = [ln.split() for ln in Path(path).read_text().splitlines() if ln.strip()]
    decanz = int(lines[0][0])
    if len(lines) < 1 + decanz:
        raise ValueError(f"decouplings.dat lists {decanz} records, has {len(lines) - 1}")
    ed = np.empty((decanz, 2), np.int32)
    es = np.empty(decanz, np.float64)
    for j in range(decanz):
        row = lines[1 + j]
        ed[j, 0], ed[j, 1] = int(row[0]), int(row[1])
        es[j] = _f_real(row[2])
    return ed, es


def decode(strnr, vnr):
    """Electrode numbers e1..e4 as decoded by bvolt (src/bvolt.f90:44-49)."""
    strnr = np.asarray(strnr, np.int64)
    vnr = np.asarray(vnr, np.int64)
    e1 = strnr % 10000
    e2 = strnr // 10000
    e3 = vnr % 10000
    e4 = vnr // 10000
    return e1, e2, e3, e4


def check_config(strnr, vnr, eanz=None):
    """Range (errnr 46) and duplicate-electrode (errnr 73) checks of rdatm."""
    e1, e2, e3, e4 = decode(strnr, vnr)
    if eanz is not None:
        for e in (e1, e2, e3, e4):
            if (e < 0).any() or (e > eanz).any():
                raise ValueError("electrode number out of range (rdatm errnr 46)")
    dup = (e1 == e2) | (e3 == e4) | (((e1 == e3) | (e1 == e4)) & (e1 != 0)) | \
        (((e2 == e3) | (e2 == e4)) & (e2 != 0))
    if dup.any():
        raise ValueError("duplicate electrodes (rdatm errnr 73)")


def write_config(path, strnr, vnr) -> None:
    body = "".join(f"{int(a)} {int(b)}\n" for a, b in zip(strnr, vnr))
    Path(path).write_text(f"{len(strnr)}\n" + body)


def read_volt(path):
    """volt.dat -> strnr, vnr, |R|, phase [mrad] (list-directed, FP32 text)."""
    lines = [ln.split() for ln in Path(path).read_text().splitlines() if ln.strip()]
    nanz = int(lines[0][0])
    arr = np.array([[float(v) for v in ln[:4]] for ln in lines[1:1 + nanz]])
    return (arr[:, 0].astype(np.int32), arr[:, 1].astype(np.int32), arr[:, 2], arr[:, 3])


def volt_fields(strnr, vnr, sigmaa):
    """The values wdatm writes (src/wdatm.f90:55-79): |R|, phase [mrad] and the
    possibly swapped vnr (potential pair swapped when round(phase/pi) != 0)."""
    sigmaa = np.asarray(sigmaa)
    bet = np.abs(sigmaa)
    pha = np.arctan2(sigmaa.imag, sigmaa.real)
    npi = np.rint(pha / math.pi) * math.pi
    swap = np.abs(npi) > 1e-12
    pha = np.where(swap, pha - npi, pha)
    ie1 = np.asarray(vnr) % 10000
    ie2 = np.asarray(vnr) // 10000
    vout = np.where(swap, ie1 * 10000 + ie2, np.asarray(vnr)).astype(np.int32)
    return bet, 1e3 * pha, vout


def list_int(i: int) -> str:
    """An integer as the reference's list-directed WRITE puts it in its files
    (gfortran: a 12-character field, right-justified)."""
    return f"{int(i):12d}"


def list_real4(x) -> str:
    """A REAL(4) value as the reference's list-directed WRITE puts it in its
    files (gfortran, src/wdatm.f90:76): a blank, then G editing with 9
    significant digits in 16 characters -- F form with four trailing blanks
    when the value, rounded to 9 digits, lies in [0.1, 1e9) (zero included),
    otherwise 1PE16.8E2."""
    v = float(np.float32(x))
    if v == 0.0:
        return " " + f"{v:.8f}".rjust(12) + "    "
    ex = int(f"{abs(v):.8e}".split("e")[1])
    if -1 <= ex < 9:
        # F editing keeps the decimal point when no digit follows it
        txt = f"{v:.{8 - ex}f}" + ("." if ex == 8 else "")
        return " " + txt.rjust(12) + "    "
    return " " + f"{v:.8E}".rjust(16)


def write_volt(path, strnr, vnr, sigmaa, full_precision=True, reference=False) -> None:
    """volt.dat: nanz, then strnr vnr |R| phase[mrad] per measurement (with
    wdatm's polarity swap).  reference=True writes the reference's own bytes
    (FP32 values in list-directed form, src/wdatm.f90:55-79); otherwise
    full_precision chooses 17 or 9 significant digits in E form."""
    bet, pha_mrad, vout = volt_fields(strnr, vnr, sigmaa)
    if reference:
        body = "".join(list_int(a) + list_int(b) + list_real4(m) + list_real4(p) + "\n"
                       for a, b, m, p in zip(strnr, vout, bet, pha_mrad))
        Path(path).write_text(list_int(len(strnr)) + "\n" + body)
        return
    fmt = "{:d} {:d} {:.17e} {:.17e}\n" if full_precision else "{:d} {:d} {:.8e} {:.8e}\n"
    if not full_precision:
        bet = bet.astype(np.float32).astype(np.float64)
        pha_mrad = pha_mrad.astype(np.float32).astype(np.float64)
    body = "".join(fmt.format(int(a), int(b), float(m), float(p))
                   for a, b, m, p in zip(strnr, vout, bet, pha_mrad))
    Path(path).write_text(f"{len(strnr)}\n" + body)


# ---- measured data (CRTomo's rdati) -----------------------------------------

class FormatError(ValueError):
    """A file the reference rejects, with the reference's error number."""

    def __init__(self, errnr: int, msg: str):
        super().__init__(f"{msg} (errnr {errnr})")
        self.errnr = errnr


class _Records:
    """List-directed reads in the manner of Fortran READ(kanal,*): each read
    starts at the next record, may continue over following records until it
    has its values, and skips the rest of its last record."""

    def __init__(self, text: str):
        self.lines = [ln.replace(",", " ").split() for ln in text.splitlines()]
        self.pos = 0

    def read(self, n: int):
        vals = []
        while len(vals) < n:
            if self.pos >= len(self.lines):
                raise EOFError
            rec = self.lines[self.pos]
            self.pos += 1
            vals.extend(rec[: n - len(vals)])
        return vals

    def backspace(self):
        self.pos = max(0, self.pos - 1)


def _f_logical(tok: str):
    t = tok.lstrip(".").lower()
    if t[:1] == "t":
        return True
    if t[:1] == "f":
        return False
    raise ValueError(tok)


def _f_int(tok: str) -> int:
    return int(tok)


def _f_real(tok: str) -> float:
    return float(tok.lower().replace("d", "e"))


@dataclasses.dataclass
class DataSet:
    """Measured data as rdati stores it (datmod/invmod arrays)."""
    strnr: np.ndarray      # (nanz,) int32, A*10^4 + B (current electrodes)
    vnr: np.ndarray        # (nanz,) int32, M*10^4 + N (potential electrodes)
    dat: np.ndarray        # (nanz,) complex128: -ln|R| - i phase[rad]
    wmatdr: np.ndarray     # 1 / sigma_{ln|R|}^2 (times the file's scale, lindiv)
    wmatdp: np.ndarray     # 1 / sigma_phase^2 (0 for DC)
    wmatd_cri: np.ndarray  # 1 / (sigma_{ln|R|}^2 + sigma_phase^2) (0 for DC)
    wdfak: np.ndarray      # (nanz,) int32, 1
    lindiv: bool           # individual errors in the file


class RefRandom:
    """The reference's pseudo-random generator (src/make_noise.f90:171-261):
    three linear congruential sequences, the first two combined into 97
    shuffled uniforms and the third choosing among them (the classic
    three-generator shuffle), and the polar method for Gaussian deviates.  The
    scale factors 1/M1 and 1/M2 are the reference's single-precision
    constants."""

    M1, IA1, IC1 = 259200, 7141, 54773
    M2, IA2, IC2 = 134456, 8121, 28441
    M3, IA3, IC3 = 243000, 4561, 51349
    N_RND = 97
    RM1 = float(np.float32(1.0) / np.float32(259200))
    RM2 = float(np.float32(1.0) / np.float32(134456))

    @staticmethod
    def _mod(a: int, p: int) -> int:  # Fortran MOD: the sign of a
        return int(math.fmod(a, p))

    def __init__(self, iseed: int = 1):
        m = self._mod
        self.ix1 = m(self.IC1 + iseed, self.M1)
        self.ix1 = m(self.IA1 * self.ix1 + self.IC1, self.M1)
        self.ix2 = m(self.ix1, self.M2)
        self.ix1 = m(self.IA1 * self.ix1 + self.IC1, self.M1)
        self.ix3 = m(self.ix1, self.M3)
        self.rnd = [0.0] * self.N_RND
        for i in range(self.N_RND):
            self.ix1 = m(self.IA1 * self.ix1 + self.IC1, self.M1)
            self.ix2 = m(self.IA2 * self.ix2 + self.IC2, self.M2)
            self.rnd[i] = (float(self.ix1) + float(self.ix2) * self.RM2) * self.RM1

    def draw(self) -> float:
        m = self._mod
        while True:
            self.ix1 = m(self.IA1 * self.ix1 + self.IC1, self.M1)
            self.ix2 = m(self.IA2 * self.ix2 + self.IC2, self.M2)
            self.ix3 = m(self.IA3 * self.ix3 + self.IC3, self.M3)
            a = self.N_RND * self.ix3  # Fortran integer division truncates toward zero
            i = 1 + (a // self.M3 if a >= 0 else -((-a) // self.M3))
            if i > self.N_RND or i < 1:
                continue
            v = self.rnd[i - 1]
            self.rnd[i - 1] = (float(self.ix1) + float(self.ix2) * self.RM2) * self.RM1
            return v

    def gauss(self) -> float:
        while True:
            x = 2.0 * self.draw() - 1.0
            y = 2.0 * self.draw() - 1.0
            z = x * x + y * y
            if z <= 1.0:
                break
        return x * math.sqrt(-2.0 * math.log(z) / z)


@dataclasses.dataclass
class NoiseModel:
    """rdati's synthetic noise (lnse; src/rall.f90:480-510): seed and the
    noise model's relative and absolute magnitude errors, phase parameters."""
    iseed: int = 1
    nstabw0: float = 0.0
    nstabm0: float = 0.0
    nstabpA1: float = 0.0
    nstabpB: float = 0.0
    nstabpA2: float = 0.0
    nstabp0: float = 0.0


def read_data(path, eanz: int, ldc: bool = False, stabw0: float = 5.0, stabm0: float = 1e-3,
              stabpA1: float = 0.0, stabpB: float = 0.0, stabpA2: float = 0.0, stabp0: float = 0.0,
              noise: NoiseModel | None = None) -> DataSet:
    """Read a measured data file (volt.dat layout) with the semantics of
    src/rdati.f90:84-363.

    * First record ``nanz [lindiv]``; without a logical second value the
      error model applies (:84-88).  The first datum's first value tells the
      CRTomo ``A*10^4+B M*10^4+N`` form (>= 10000) from ``A B M N`` (:90-94).
    * Individual errors (lindiv): each datum carries the magnitude error in
      Ohm and, unless DC, the phase error in mrad; a last record holds global
      factors applied to the weights (:144-214, :344-363).
    * Error model: sigma_{ln|R|} = stabw0/100 + stabm0/|R| and sigma_phase =
      (stabpA1 |R|^stabpB + stabpA2/100 |phase| + stabp0) mrad (:216-254).
    * Errors as the reference numbers them: 73 duplicate electrodes (lindiv
      only, as there), 46 electrode out of range, 88 non-positive error, 94
      non-positive |R|, 95 |phase| > 1000 pi mrad, 2 premature end of file,
      3 unreadable value.
    * ``noise`` adds the reference's synthetic noise (lnse, :110-135 and
      :256-292): Gaussian deviates from the reference's generator seeded with
      iseed (magnitudes) and -iseed (phases), scaled by eps_R = nstabw0/100 |R| +
      nstabm0 and eps_phi = nstabpA1 eps_R^nstabpB + nstabpA2/100 |phase| +
      nstabp0, added to the data of error-model files after their weights
      are formed.  The reference's listing files are not written.
    """
    pi = math.acos(-1.0)
    rec = _Records(Path(path).read_text())
    try:
        first = rec.read(1)
        nanz = _f_int(first[0])
        lindiv = False
        line0 = rec.lines[0]
        if len(line0) >= 2:
            try:
                lindiv = _f_logical(line0[1])
            except ValueError:
                lindiv = False
        probe = rec.read(1)
        rec.backspace()
        crtf = _f_int(probe[0]) - 10000 >= 0
        strnr = np.zeros(nanz, np.int32)
        vnr = np.zeros(nanz, np.int32)
        dat = np.zeros(nanz, np.complex128)
        wmatdr = np.zeros(nanz)
        wmatdp = np.zeros(nanz)
        wcri = np.zeros(nanz)
        ncol = (2 if crtf else 4) + (1 if ldc else 2) + ((1 if ldc else 2) if lindiv else 0)
        pha = 0.0
        if noise is not None:
            g = RefRandom(noise.iseed)
            rnd_r = [g.gauss() for _ in range(nanz)]
            if not ldc:
                g = RefRandom(-noise.iseed)
                rnd_p = [g.gauss() for _ in range(nanz)]
        for i in range(nanz):
            v = rec.read(ncol)
            if crtf:
                strnr[i], vnr[i] = _f_int(v[0]), _f_int(v[1])
                k = 2
            else:
                e1, e2, e3, e4 = (_f_int(x) for x in v[:4])
                strnr[i], vnr[i] = e1 * 10000 + e2, e3 * 10000 + e4
                k = 4
            bet = _f_real(v[k])
            if not ldc:
                pha = _f_real(v[k + 1])
            stabwp = 0.0
            if lindiv:
                stabw = _f_real(v[k + (1 if ldc else 2)])
                if not ldc:
                    stabwp = _f_real(v[k + 3])
                    if stabwp <= 0.0:
                        raise FormatError(88, f"non-positive phase error of datum {i + 1}")
                    stabwp = 1e-3 * stabwp
                with np.errstate(divide="ignore", invalid="ignore"):
                    stabw = float(np.float64(stabw) / np.float64(bet))
                e1, e2 = int(strnr[i]) % 10000, int(strnr[i]) // 10000
                e3, e4 = int(vnr[i]) % 10000, int(vnr[i]) // 10000
                if (e1 == e2 or e3 == e4 or ((e1 == e3 or e1 == e4) and e1 != 0)
                        or ((e2 == e3 or e2 == e4) and e2 != 0)):
                    raise FormatError(73, f"duplicate electrodes for reading {i + 1}")
                if min(e1, e2, e3, e4) < 0 or max(e1, e2, e3, e4) > eanz:
                    raise FormatError(46, f"electrode pair {i + 1} not correct")
                if stabw <= 0.0:
                    raise FormatError(88, f"non-positive magnitude error of datum {i + 1}")
            else:
                if not ldc:
                    stabwp = (stabpA1 * bet ** stabpB + 1e-2 * stabpA2 * abs(pha) + stabp0) * 1e-3
                if bet <= 0.0:
                    raise FormatError(94, f"non-positive resistance of datum {i + 1}")
                stabw = 1e-2 * stabw0 + stabm0 / bet
                if noise is not None:
                    eps_r = 1e-2 * noise.nstabw0 * bet + noise.nstabm0
                    if not ldc:
                        eps_p = (noise.nstabpA1 * eps_r ** noise.nstabpB + 1e-2 * noise.nstabpA2 * abs(pha)
                                 + noise.nstabp0)
                        pha = pha + rnd_p[i] * eps_p
                    bet = bet + rnd_r[i] * eps_r
            if bet <= 0.0:
                raise FormatError(94, f"non-positive resistance of datum {i + 1}")
            if ldc:
                pha = 0.0
            elif abs(pha) > 1e3 * pi:
                raise FormatError(95, f"phase of datum {i + 1} beyond 1000 pi mrad")
            dat[i] = complex(-math.log(bet), -pha / 1e3)
            wmatdr[i] = 1.0 / (stabw ** 2.0)
            if not ldc:
                wcri[i] = 1.0 / (stabw ** 2.0 + stabwp ** 2.0)
                wmatdp[i] = 1.0 / (stabwp ** 2.0)
            e1, e2 = int(strnr[i]) % 10000, int(strnr[i]) // 10000
            e3, e4 = int(vnr[i]) % 10000, int(vnr[i]) // 10000
            if min(e1, e2, e3, e4) < 0 or max(e1, e2, e3, e4) > eanz:
                raise FormatError(46, f"electrode pair {i + 1} not correct")
        if lindiv:
            if ldc:
                s = _f_real(rec.read(1)[0])
                if s <= 0.0:
                    raise FormatError(88, "non-positive global error factor")
                wmatdr = wmatdr * s * s
            else:
                s, sp = (_f_real(x) for x in rec.read(2))
                if s <= 0.0 or sp <= 0.0:
                    raise FormatError(88, "non-positive global error factor")
                wmatdr = wmatdr * s * s
                wmatdp = wmatdp * sp * sp
    except EOFError:
        raise FormatError(2, "premature end of the data file") from None
    except (ValueError, IndexError) as e:
        if isinstance(e, FormatError):
            raise
        raise FormatError(3, f"unreadable value in the data file: {e}") from None
    return DataSet(strnr, vnr, dat, wmatdr, wmatdp, wcri, np.ones(nanz, np.int32), lindiv)
