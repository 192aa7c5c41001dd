"""Seeded synthetic tomodirs for the benchmark workloads.

Every generator returns a :class:`~titatest_amd.formats.Mesh` that the reference's
own ``relem`` would accept: area elements first, then boundary types, boundary
elements oriented clockwise so the beta normal points outwards
(src/beta.f90:93-96), ``rnr`` linking each boundary element to its area element,
natural column node ordering and ``mb`` = the largest node difference inside an
element (what CutMcK would write for these structured grids).

Workloads (sizes follow the reference rules for wavenumbers and dipole-dipole
counting):

* ``small_block``: 40x20 quads of 1 m x 0.5 m (861 nodes), 20 surface electrodes at
  1 m, 100 Ohm.m / -5 mrad with a 4x3 block of 10 Ohm.m / -20 mrad at 3 m depth,
  dipole-dipole a=1, n=1..8 (108 quadrupoles), mixed boundaries.
* ``large_grf``: 240x80 quads (19,521 nodes); 96 uniform 1.25 m cells under the
  electrodes, widths growing x1.1 outside; 96 electrodes; log10(rho) and phase
  from seeded Gaussian random fields; dipole-dipole a=1..4, n=1..10 (3,090).
* ``uneven_batch``: 64 models, nx~U[30,160], nz~U[15,50], E~U[24,72].
* ``topo_triangles``: terrain-following triangles, z_top = 2 sin(2 pi x / 40).
"""
from __future__ import annotations

import math

import numpy as np

from .formats import Mesh


def _quad_mesh(xn: np.ndarray, zn: np.ndarray, side_type: int = 11) -> Mesh:
    """Quadrilateral type-8 mesh on the node lines xn (left->right) and zn (top
    surface first, increasingly negative).  Node numbering runs down each column
    (node = ix*(nz+1) + iz + 1), so mb = nz + 2.  The top is a no-flow (type 12)
    boundary; left, right and bottom are of ``side_type`` (11 mixed, 13 Dirichlet)."""
    nx = len(xn) - 1
    nz = len(zn) - 1
    ncol = nz + 1
    sanz = (nx + 1) * ncol
    X, Z = np.meshgrid(xn, zn, indexing="ij")
    sx = X.reshape(-1).astype(np.float64)
    sy = Z.reshape(-1).astype(np.float64)

    def node(ix, iz):
        return ix * ncol + iz + 1

    # elements row by row from the surface down (k = (iz)*nx + ix + 1), the order
    # the rectangular smoothness operator assumes (src/bsmatm_mod.f90:222)
    iz, ix = np.meshgrid(np.arange(nz), np.arange(nx), indexing="ij")
    ix = ix.reshape(-1)
    iz = iz.reshape(-1)
    # counter-clockwise in (x, y) with y up: lower-left, lower-right, upper-right, upper-left
    quads = np.stack([node(ix, iz + 1), node(ix + 1, iz + 1), node(ix + 1, iz), node(ix, iz)], 1)
    elem_of = lambda a, b: b * nx + a + 1  # noqa: E731  (1-based element of cell (ix=a, iz=b))

    top = [(node(i, 0), node(i + 1, 0), elem_of(i, 0)) for i in range(nx)]
    # clockwise: right side downwards, bottom right->left, left side upwards
    right = [(node(nx, j), node(nx, j + 1), elem_of(nx - 1, j)) for j in range(nz)]
    bottom = [(node(i + 1, nz), node(i, nz), elem_of(i, nz - 1)) for i in range(nx - 1, -1, -1)]
    left = [(node(0, j + 1), node(0, j), elem_of(0, j)) for j in range(nz - 1, -1, -1)]
    mixed_b = right + bottom + left
    typ = np.array([8, 12, side_type], np.int32)
    bnd = [top, mixed_b]
    nel = nx * nz + sum(len(b) for b in bnd)
    nrel = np.zeros((nel, 4), np.int32)
    nrel[: nx * nz] = quads
    rnr = []
    p = nx * nz
    for b in bnd:
        for n1, n2, e in b:
            nrel[p, 0], nrel[p, 1] = n1, n2
            rnr.append(e)
            p += 1
    nelanz = np.array([nx * nz] + [len(b) for b in bnd], np.int32)
    selanz = np.array([4] + [2] * len(bnd), np.int32)
    mesh = Mesh(sanz=sanz, mb=nz + 2, typ=typ, nelanz=nelanz, selanz=selanz,
                snr=np.arange(1, sanz + 1, dtype=np.int32), sx=sx, sy=sy, nrel=nrel,
                rnr=np.array(rnr, np.int32))
    return mesh


def surface_nodes(mesh_nz: int, ix) -> np.ndarray:
    """1-based node numbers of surface nodes in columns ix (column ordering)."""
    return (np.asarray(ix) * (mesh_nz + 1) + 1).astype(np.int32)


def dipole_dipole(eanz: int, a_values, n_values):
    """strnr/vnr for dipole-dipole: A=i, B=i+a, N=B+n*a, M=N+a (all <= eanz).

    Encoded as strnr = A*10^4 + B, vnr = M*10^4 + N, as in the reference's
    fixtures (e.g. tests/minimal/Reference/mod/volt.dat: 10002 40003)."""
    st, vn = [], []
    for a in a_values:
        for n in n_values:
            for i in range(1, eanz + 1):
                A, B = i, i + a
                N = B + n * a
                M = N + a
                if M > eanz:
                    break
                st.append(A * 10000 + B)
                vn.append(M * 10000 + N)
    return np.array(st, np.int32), np.array(vn, np.int32)


def grf_on_points(px, pz, rng, mean, std, corr_range, h=0.5, pad=None):
    """Stationary Gaussian random field with exponential covariance
    C(r) = std^2 exp(-r / corr_range), sampled at points (px, pz) by
    FFT circulant embedding on a regular h-spaced grid, then nearest-cell lookup.
    Points outside the grid are clamped to its edge."""
    pad = corr_range * 3 if pad is None else pad
    x0, x1 = px.min(), px.max()
    z0, z1 = pz.min(), pz.max()
    nxg = int(math.ceil((x1 - x0) / h)) + 1
    nzg = int(math.ceil((z1 - z0) / h)) + 1
    mx = int(2 ** math.ceil(math.log2(nxg + pad / h)))
    mz = int(2 ** math.ceil(math.log2(nzg + pad / h)))
    dx = np.minimum(np.arange(mx), mx - np.arange(mx)) * h
    dz = np.minimum(np.arange(mz), mz - np.arange(mz)) * h
    r = np.sqrt(dx[:, None] ** 2 + dz[None, :] ** 2)
    cov = np.exp(-r / corr_range)
    lam = np.fft.fft2(cov).real
    lam = np.maximum(lam, 0.0)
    noise = rng.standard_normal((mx, mz)) + 1j * rng.standard_normal((mx, mz))
    field = np.fft.fft2(np.sqrt(lam / (mx * mz)) * noise).real[:nxg, :nzg]
    gi = np.clip(np.rint((px - x0) / h).astype(int), 0, nxg - 1)
    gk = np.clip(np.rint((pz - z0) / h).astype(int), 0, nzg - 1)
    return mean + std * field[gi, gk]


def centroids(mesh: Mesh):
    """Area-element centroids (x, z)."""
    elanz = mesh.elanz
    cx = np.zeros(elanz)
    cz = np.zeros(elanz)
    for t, nk, a, b in mesh.element_ranges():
        if t > 10:
            continue
        x, z = mesh.node_xy(mesh.nrel[a:b, :nk])
        cx[a:b] = x.mean(1)
        cz[a:b] = z.mean(1)
    return cx, cz


def sigma_from_rho(mag, pha_mrad):
    """sigma = e^{-i phi} / |rho| (src/rsigma.f90:187)."""
    pha = 1e-3 * np.asarray(pha_mrad)
    return np.cos(pha) / mag - 1j * np.sin(pha) / mag


class Workload:
    """A synthetic tomodir: mesh + electrodes + model + measurement configurations."""

    def __init__(self, name, mesh, mag, pha_mrad, strnr, vnr):
        self.name = name
        self.mesh = mesh
        self.mag = np.asarray(mag, np.float64)
        self.pha_mrad = np.asarray(pha_mrad, np.float64)
        self.sigma = sigma_from_rho(self.mag, self.pha_mrad)
        self.strnr = strnr
        self.vnr = vnr

    @property
    def nanz(self):
        return int(len(self.strnr))

    def write_tomodir(self, root):
        from pathlib import Path
        from . import formats as F
        root = Path(root)
        for sub in ("grid", "rho", "config", "mod"):
            (root / sub).mkdir(parents=True, exist_ok=True)
        F.write_elem(root / "grid" / "elem.dat", self.mesh)
        F.write_elec(root / "grid" / "elec.dat", self.mesh.enr)
        F.write_rho(root / "rho" / "rho.dat", self.mag, self.pha_mrad)
        F.write_config(root / "config" / "config.dat", self.strnr, self.vnr)


def small_block() -> Workload:
    """40x20 quads of 1 m x 0.5 m, 20 electrodes at x = 10..29 m, block anomaly."""
    xn = np.arange(41, dtype=np.float64)
    zn = -0.5 * np.arange(21, dtype=np.float64)
    mesh = _quad_mesh(xn, zn)
    mesh.enr = surface_nodes(20, np.arange(10, 30))
    cx, cz = centroids(mesh)
    mag = np.full(mesh.elanz, 100.0)
    pha = np.full(mesh.elanz, -5.0)
    blk = (cx > 18) & (cx < 22) & (cz < -3.0) & (cz > -4.5)
    mag[blk] = 10.0
    pha[blk] = -20.0
    st, vn = dipole_dipole(20, [1], range(1, 9))
    return Workload("small_block", mesh, mag, pha, st, vn)


def small_dirichlet() -> Workload:
    """As small_block but with homogeneous Dirichlet (type 13) sides and bottom:
    no mixed boundaries, so one factor per wavenumber serves every electrode."""
    w = small_block()
    xn = np.arange(41, dtype=np.float64)
    zn = -0.5 * np.arange(21, dtype=np.float64)
    mesh = _quad_mesh(xn, zn, side_type=13)
    mesh.enr = w.mesh.enr
    return Workload("small_dirichlet", mesh, w.mag, w.pha_mrad, w.strnr, w.vnr)


def boundary_values(w: Workload, every=4, seed=7, inset=0):
    """Prescribed potentials for randb2 (the rwdnr/rwd arrays rrandb reads,
    src/rrandb.f90): every ``every``-th node of the bottom boundary and the
    three lowest nodes of the left side of a column-numbered quad mesh (moved
    ``inset`` node rows up and columns right), with complex values of the
    order of the source potentials there."""
    m = w.mesh
    ncol = int(round(m.sanz / (len(np.unique(m.sx)))))   # nodes per column
    nxn = m.sanz // ncol
    bottom = [ix * ncol + ncol - inset for ix in range(inset, nxn - inset, every)]
    left = [inset * ncol + ncol - 1 - inset - j for j in range(1, 4)]
    rwdnr = np.array(bottom + left, np.int32)
    rng = np.random.default_rng(seed)
    rwd = 0.05 * rng.uniform(0.5, 1.5, len(rwdnr)) * np.exp(1j * rng.uniform(-0.3, 0.3, len(rwdnr)))
    return rwdnr, rwd.astype(np.complex128)


def _graded_axis(n_uniform, dx, n_side, factor):
    """Node coordinates: n_uniform cells of width dx, n_side cells growing by
    `factor` on each side."""
    w = dx * factor ** np.arange(1, n_side + 1)
    left = -np.cumsum(w)[::-1]
    mid = dx * np.arange(n_uniform + 1)
    right = mid[-1] + np.cumsum(w)
    return np.concatenate([left, mid, right])


def _depth_axis(nz, dz0, factor):
    return -np.concatenate([[0.0], np.cumsum(dz0 * factor ** np.arange(nz))])


def grf_model(mesh, seed_mag, seed_pha):
    """log10(rho) ~ GRF(mean 2, sd 0.5, exp. cov., range 10 m); phase from an
    independent field mapped affinely onto [-30, -2] mrad."""
    cx, cz = centroids(mesh)
    # the random fields are generated over the central, finely resolved region;
    # cells outside take the nearest value.
    core = (np.abs(cx - np.median(cx)) < 150) & (cz > -80)
    x_lo, x_hi = cx[core].min(), cx[core].max()
    z_lo = cz[core].min()
    px = np.clip(cx, x_lo, x_hi)
    pz = np.clip(cz, z_lo, 0.0)
    lmag = grf_on_points(px, pz, np.random.Generator(np.random.PCG64(seed_mag)), 2.0, 0.5, 10.0)
    g = grf_on_points(px, pz, np.random.Generator(np.random.PCG64(seed_pha)), 0.0, 1.0, 10.0)
    g = (g - g.min()) / max(g.max() - g.min(), 1e-30)
    pha = -2.0 - 28.0 * g
    return 10.0 ** lmag, pha


def large_grf(nx_side=72, n_uniform=96, nz=80, eanz=96, seed=1001, side_type=11) -> Workload:
    """240x80 quads, 96 electrodes at 1.25 m, GRF model, dipole-dipole a=1..4, n=1..10.
    side_type 13 gives Dirichlet sides and bottom (one factor per wavenumber)."""
    xn = _graded_axis(n_uniform, 1.25, nx_side, 1.1)
    zn = _depth_axis(nz, 0.3125, 1.05)
    mesh = _quad_mesh(xn, zn, side_type=side_type)
    mesh.enr = surface_nodes(nz, nx_side + np.arange(eanz))
    mag, pha = grf_model(mesh, seed, seed + 1)
    st, vn = dipole_dipole(eanz, range(1, 5), range(1, 11))
    return Workload("large_grf", mesh, mag, pha, st, vn)


def uneven_batch(count=64, seed=1006):
    """`count` independent models with nx~U[30,160], nz~U[15,50], E~U[24,72]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for m in range(count):
        nx = int(rng.integers(30, 161))
        nz = int(rng.integers(15, 51))
        eanz = int(rng.integers(24, 73))
        eanz = min(eanz, nx - 1)
        side = (nx - eanz) // 2
        n_uniform = nx - 2 * side
        xn = _graded_axis(n_uniform, 1.25, side, 1.1)
        zn = _depth_axis(nz, 0.3125, 1.05)
        mesh = _quad_mesh(xn, zn)
        mesh.enr = surface_nodes(nz, side + np.arange(eanz))
        mag, pha = grf_model(mesh, seed + 10 + 2 * m, seed + 11 + 2 * m)
        st, vn = dipole_dipole(eanz, range(1, 5), range(1, 11))
        out.append(Workload(f"uneven_{m}", mesh, mag, pha, st, vn))
    return out


def topo_triangles(nx_nodes=257, nz_nodes=100, eanz=128, length=160.0, depth=60.0,
                   seed=1011) -> Workload:
    """Terrain-following triangular mesh (type 3), z_top = 2 sin(2 pi x / 40).

    Each structured cell is split into two counter-clockwise triangles; node
    numbering is column-wise, so mb = nz_nodes + 1.  The default 257 x 100
    nodes (0.625 m columns over 160 m) put the 128 electrodes on every second
    surface node, 1.25 m apart."""
    xn = np.linspace(0.0, length, nx_nodes)
    s = np.linspace(0.0, 1.0, nz_nodes) ** 1.5
    ncol = nz_nodes
    sanz = nx_nodes * ncol
    ztop = 2.0 * np.sin(2.0 * np.pi * xn / 40.0)
    sx = np.repeat(xn, ncol)
    sy = (ztop[:, None] - s[None, :] * (depth + ztop[:, None])).reshape(-1)

    def node(ix, iz):
        return ix * ncol + iz + 1

    tris = []
    cell = {}
    for ix in range(nx_nodes - 1):
        for iz in range(nz_nodes - 1):
            a, b, c, d = node(ix, iz + 1), node(ix + 1, iz + 1), node(ix + 1, iz), node(ix, iz)
            cell[(ix, iz)] = len(tris)
            tris.append((a, b, c))     # lower-left, lower-right, upper-right
            tris.append((a, c, d))     # lower-left, upper-right, upper-left
    nxe, nze = nx_nodes - 1, nz_nodes - 1
    top = [(node(i, 0), node(i + 1, 0), cell[(i, 0)] + 2) for i in range(nxe)]
    right = [(node(nxe, j), node(nxe, j + 1), cell[(nxe - 1, j)] + 1) for j in range(nze)]
    bottom = [(node(i + 1, nze), node(i, nze), cell[(i, nze - 1)] + 1) for i in range(nxe - 1, -1, -1)]
    left = [(node(0, j + 1), node(0, j), cell[(0, j)] + 2) for j in range(nze - 1, -1, -1)]
    mixed_b = right + bottom + left
    nel = len(tris) + len(top) + len(mixed_b)
    nrel = np.zeros((nel, 3), np.int32)
    nrel[: len(tris)] = np.array(tris, np.int32)
    rnr = []
    p = len(tris)
    for n1, n2, e in top + mixed_b:
        nrel[p, 0], nrel[p, 1] = n1, n2
        rnr.append(e)
        p += 1
    mesh = Mesh(sanz=sanz, mb=ncol + 1, typ=np.array([3, 12, 11], np.int32),
                nelanz=np.array([len(tris), len(top), len(mixed_b)], np.int32),
                selanz=np.array([3, 2, 2], np.int32), snr=np.arange(1, sanz + 1, dtype=np.int32),
                sx=sx, sy=sy, nrel=nrel, rnr=np.array(rnr, np.int32))
    # electrodes: 128 surface nodes at ~1.25 m spacing, centred on the line
    spacing = xn[1] - xn[0]
    step = max(1, int(round(1.25 / spacing)))
    first = (nx_nodes - step * (eanz - 1)) // 2
    if first < 0:
        raise ValueError(f"{eanz} electrodes every {step} columns do not fit on {nx_nodes} node columns")
    cols = first + step * np.arange(eanz)
    mesh.enr = (cols * ncol + 1).astype(np.int32)
    mag, pha = grf_model(mesh, seed, seed + 1)
    st, vn = dipole_dipole(eanz, range(1, 7), range(1, 13))
    return Workload("topo_triangles", mesh, mag, pha, st, vn)
