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
  potential pair when the phase is near +-pi) and src/rdati.f90:84-240 (reader).

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


def write_volt(path, strnr, vnr, sigmaa, full_precision=True) -> None:
    bet, pha_mrad, vout = volt_fields(strnr, vnr, sigmaa)
    fmt = "{:d} {:d} {:.17e} {:.17e}\n" if full_precision else "{:d} {:d} {:.8e} {:.8e}\n"
    if not full_precision:
        bet = bet.astype(np.float32).astype(np.float64)
        pha_mrad = pha_mrad.astype(np.float32).astype(np.float64)
    body = "".join(fmt.format(int(a), int(b), float(m), float(p))
                   for a, b, m, p in zip(strnr, vout, bet, pha_mrad))
    Path(path).write_text(f"{len(strnr)}\n" + body)
