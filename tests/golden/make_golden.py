"""Write full-precision golden vectors from the reference's own routines.

Run in a container that has the reference tree (for ``make -C oracle ref``):

    make -C oracle ref && python tests/golden/make_golden.py

It drives oracle/_ref/libcrtref.so (the reference's Fortran routines compiled
from their source, plus oracle/ref_driver.f90) on small cases and stores inputs
that are not regenerated by the tests plus the reference's outputs in
``tests/golden/ref_<case>.npz``.  The tests compare the C oracle and the HIP path
against these files; neither the reference nor its build is needed to run them.
"""
from __future__ import annotations

import sys
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[2]
sys.path.insert(0, str(ROOT))

from oracle.ref import Reference, available  # noqa: E402
from titatest_amd import formats as F  # noqa: E402
from titatest_amd import meshgen  # noqa: E402

OUT = Path(__file__).resolve().parent
FIX = OUT / "fixtures"

# rows of the sensitivity matrices kept (all columns)
SENS_ROWS = 8


def electrode_slice(kpot, enr):
    """kpot at the electrode nodes: (E_nodes, E_sources, K)."""
    return kpot[np.asarray(enr) - 1, :, :]


def synthetic(w, lsytop=False, nelbg=64):
    m = w.mesh
    r = Reference(m, lsytop=lsytop)
    r.load_data(w.strnr, w.vnr)
    kwn, kwnwi = r.rwaven()
    elbg, relbg, kg = r.precal()
    kpot, hpot = r.forward(w.sigma, nthreads=8)
    v0, sa0 = r.volt(0)
    s0 = r.sens(0)
    v1, sa1 = r.volt(1)
    s1 = r.sens(1)
    nimn = m.maxnodes * (m.maxnodes + 1) // 2
    K = len(kwn)
    elbg = elbg.reshape((m.elanz, nimn, K), order="F")
    relanz = m.relanz
    kg = kg.reshape((max(relanz, 1), m.eanz, K), order="F")
    relbg = relbg.reshape((max(relanz, 1), nimn), order="F")
    btyp = np.concatenate([np.full(n, t) for t, n in zip(m.typ, m.nelanz) if t > 10])
    mixed = btyp == 11
    np.savez_compressed(
        OUT / f"ref_{w.name}.npz",
        kwn=kwn, kwnwi=kwnwi, lbeta=r.lbeta, lrandb=r.lrandb, sytop=r.sytop,
        elbg=elbg[:nelbg], relbg=relbg[mixed][:, :3], kg=kg[mixed][:64],
        volt=v0, sigmaa0=sa0, sigmaa1=sa1,
        hpot_elec=hpot[np.asarray(m.enr) - 1, :],
        kpot_elec=electrode_slice(kpot, m.enr),
        kpot_col=kpot[:, 0, :],
        sens0=s0[:SENS_ROWS], sens1=s1[:SENS_ROWS],
    )
    print(f"wrote ref_{w.name}.npz  K={K} N={len(v0)}")


def synthetic_dc(w):
    """The DC twins (kompadc ... bsendc) on the same case: kpotdc at the
    electrode nodes, hpotdc electrode rows, DC voltages and sensdc rows, plus
    one scaled system and its choldc factor."""
    m = w.mesh
    r = Reference(m)
    r.load_data(w.strnr, w.vnr)
    r.rwaven()
    r.precal()
    kp, hp = r.forward_dc(w.sigma, nthreads=8)
    v1, sa1 = r.volt(1)
    sdc = r.sens_dc()
    a, b, fak = r.assemble_dc(w.sigma, 0, 0)
    L = r.choldc(a)
    np.savez_compressed(
        OUT / f"ref_{w.name}_dc.npz",
        kpotdc_elec=electrode_slice(kp, m.enr), kpotdc_col=kp[:, 0, :],
        hpotdc_elec=hp[np.asarray(m.enr) - 1, :], volt=v1, sigmaa1=sa1,
        sensdc=sdc[:SENS_ROWS], sensdc_rowsum=sdc.sum(axis=1),
        a00=a, b00=b, fak00=fak, chol00=L,
    )
    print(f"wrote ref_{w.name}_dc.npz  N={len(v1)}")


def synthetic_rb2(w, inset):
    """Prescribed boundary potentials (randb2 / randbdc2) on the same case,
    values from meshgen.boundary_values(w, inset=inset)."""
    m = w.mesh
    rn, rv = meshgen.boundary_values(w, inset=inset)
    r = Reference(m)
    r.load_data(w.strnr, w.vnr)
    r.rwaven()
    r.precal()
    r.set_boundary_values(rn, rv)
    a, b, fak = r.assemble(w.sigma, 0, 0)
    kpot, hpot = r.forward(w.sigma, nthreads=8)
    v1, sa1 = r.volt(1)
    kd, hd = r.forward_dc(w.sigma, nthreads=8)
    vd, _ = r.volt(1)
    r.set_boundary_values([], [])
    np.savez_compressed(
        OUT / f"ref_{w.name}_rb2.npz", rwdnr=rn, rwd=rv, a00=a, b00=b, fak00=fak,
        kpot_elec=electrode_slice(kpot, m.enr), kpot_col=kpot[:, 0, :],
        hpot_elec=hpot[np.asarray(m.enr) - 1, :], volt=v1, sigmaa1=sa1,
        kpotdc_elec=electrode_slice(kd, m.enr), hpotdc_elec=hd[np.asarray(m.enr) - 1, :], volt_dc=vd,
    )
    print(f"wrote ref_{w.name}_rb2.npz  N={len(v1)}")


def triangular_smoothness():
    """bnachbar + bsmatmtri (ltri = 1) on a small terrain-following
    triangular mesh, alfx = 1, alfz = 0.5."""
    w = meshgen.topo_triangles(nx_nodes=25, nz_nodes=10, eanz=12)
    r = Reference(w.mesh)
    sm, nb = r.bsmatmtri(1.0, 0.5)
    np.savez_compressed(OUT / "ref_tri_small_smooth.npz", smatm=sm, nachbar=nb)
    print(f"wrote ref_tri_small_smooth.npz  manz={sm.shape[0]}")


def topo_small():
    """A small terrain-following triangular mesh (elem3 on sheared triangles,
    30 m of the 2 sin(2 pi x / 40) surface so the no-flow mean is not zero)
    with lsytop: the mixed-boundary image sources mirrored about bsytop's
    height (src/bsytop.f90:35-69, src/beta.f90:78-90).  Every element matrix
    is stored."""
    w = meshgen.topo_triangles(nx_nodes=31, nz_nodes=12, eanz=16, length=30.0, depth=12.0, seed=1013)
    w.name = "topo_tri_small"
    synthetic(w, lsytop=True, nelbg=w.mesh.elanz)


# (l, k) systems whose full kpot column the triangular fixture stores
TRI_FIXTURE_SYSTEMS = ((0, 0), (107, 5), (214, 10))


def fixture_triangular(name="modelling_triangular"):
    """The reference's own triangular modelling case (tests/modelling_triangular:
    1,489 nodes, band width 199, 215 electrodes, 90 measurements from its
    config.dat), run through the reference routines: voltages both modes,
    electrode rows of hpot, three full kpot columns, the first sensitivity
    rows."""
    d = FIX / name
    m = F.read_elem(d / "elem.dat")
    m.enr = F.read_elec(d / "elec.dat")
    sig = F.read_rho(d / "rho.dat", m.elanz)
    st, vn = F.read_config(d / "config.dat", m.eanz)
    r = Reference(m)
    r.load_data(st, vn)
    kwn, kwnwi = r.rwaven()
    r.precal()
    kpot, hpot = r.forward(sig, nthreads=8)
    v0, sa0 = r.volt(0)
    v1, sa1 = r.volt(1)
    s1 = r.sens(1)
    sysl = np.array([s for s in TRI_FIXTURE_SYSTEMS if s[1] < len(kwn)], np.int32)
    np.savez_compressed(
        OUT / f"ref_{name}.npz", kwn=kwn, kwnwi=kwnwi, volt=v0, sigmaa0=sa0, sigmaa1=sa1,
        hpot_elec=hpot[np.asarray(m.enr) - 1, :], systems=sysl,
        kpot_sys=np.stack([kpot[:, l, k] for l, k in sysl]), sens1=s1[:4])
    print(f"wrote ref_{name}.npz  K={len(kwn)} N={len(v0)} mb={m.mb}")


def smoothness_cases():
    """bsmatmreg (ltri = 0) on a graded quadrilateral mesh (geometric growth
    sideways and with depth, so every weight has its own cell geometry), and
    bsmatmtri with decoupled edges: the reference's own decouplings.dat files
    on their grid (tests/GridDecoupling/regular_grids), and seeded pairs on a
    terrain-following triangular mesh (one pair listed twice, one reversed)."""
    w = meshgen.large_grf(nx_side=5, n_uniform=14, nz=9, eanz=12, seed=1021)
    nx = 14 + 2 * 5
    r = Reference(w.mesh)
    out = dict(reg_nx=nx, reg_nz=9)
    out["reg_smatm"] = r.bsmatmreg(nx, 9, 1.0, 0.5)
    out["reg_smatm_b"] = r.bsmatmreg(nx, 9, 3.0, 0.25)
    d = FIX / "decoupling"
    m = F.read_elem(d / "elem.dat")
    m.enr = F.read_elec(d / "elec.dat")
    r = Reference(m)
    for tag in ("line", "full"):
        ed, es = F.read_decouplings(d / f"decouplings_{tag}.dat")
        sm, nb = r.bsmatmtri(1.0, 1.0, (ed, es))
        out[f"dec_{tag}_smatm"] = sm
    out["dec_nachbar"] = nb
    wt = meshgen.topo_triangles(nx_nodes=25, nz_nodes=10, eanz=12)
    r = Reference(wt.mesh)
    sm0, nb = r.bsmatmtri(1.0, 0.5)
    rng = np.random.default_rng(1022)
    el = rng.choice(wt.mesh.elanz, 24, replace=False)
    pairs = []
    for e in el:
        nbs = [x for x in nb[e, :3] if x > 0]
        pairs.append((e + 1, int(rng.choice(nbs))))
    pairs.append(pairs[0])                       # the same edge twice: two factors
    pairs.append((pairs[1][1], pairs[1][0]))     # reversed orientation
    ed = np.array(pairs, np.int32)
    es = rng.uniform(0.0, 1.0, len(pairs))
    es[3] = 0.0
    sm, _ = r.bsmatmtri(1.0, 0.5, (ed, es))
    out.update(tri_edecoup=ed, tri_edecstr=es, tri_smatm=sm)
    np.savez_compressed(OUT / "ref_smoothness.npz", **out)
    print("wrote ref_smoothness.npz")


def fixture(name, swrtr, lsink=False, nsink=0):
    d = FIX / name
    m = F.read_elem(d / "elem.dat")
    m.enr = F.read_elec(d / "elec.dat")
    sig = F.read_rho(d / "rho.dat", m.elanz)
    st, vn, mag, pha = F.read_volt(d / "volt.dat")
    r = Reference(m, lsink=lsink, nsink=nsink)
    r.load_data(st, vn)
    if swrtr:
        r.rwaven()
    else:
        r.set_waven([0.0], [1.0], swrtr=0)
    r.precal()
    kpot, hpot = r.forward(sig, nthreads=8)
    v0, sa0 = r.volt(0)
    np.savez_compressed(OUT / f"ref_{name}.npz", volt=v0, sigmaa0=sa0,
                        hpot_elec=hpot[np.asarray(m.enr) - 1, :])
    print(f"wrote ref_{name}.npz  N={len(v0)}")


def inversion_steps():
    """chkpo2 over a grid of datum/response phases spanning its range
    classes (both ldum), dmisft on seeded data under every switch combination
    (lpol, lfpi, ldiff, lrobust), and bpar / bsigma on the small block mesh
    with an identity and a zoned element-to-parameter map."""
    from oracle import ref
    pi = np.pi
    ph = np.array([0.0, 0.1, -0.1, 1.5, -1.5, pi / 2, -pi / 2, np.nextafter(pi / 2, 4.0),
                   np.nextafter(-pi / 2, -4.0), 2.0, -2.0, 3.1, -3.1])
    cin = []
    cout = []
    for a in ph:
        for b in ph:
            for ld in (0, 1):
                d, s = complex(1.3, a), complex(-0.7, b)
                cd, cs, wd = ref.chkpo2(d, s, ld)
                cin.append((d, s, ld))
                cout.append((cd, cs, wd))
    rng = np.random.default_rng(2024)
    N = 400
    dat = rng.normal(0, 1, N) + 1j * rng.uniform(-2.5, 2.5, N)
    sig = dat + 0.05 * (rng.normal(size=N) + 1j * rng.normal(size=N))
    sig[::17] += 2.0j
    wmatd = 1.0 / (0.02 + 1e-3 * rng.uniform(size=N)) ** 2
    wmatdr = rng.uniform(0.5, 2, N) * wmatd
    wmatdp = rng.uniform(0.5, 2, N) * wmatd
    d0 = rng.normal(size=N) + 0.01j * rng.normal(size=N)
    fm0 = d0 + 0.01 * rng.normal(size=N)
    flag_sets = np.arange(16, dtype=np.int32)
    l1min = 1.0
    stats = np.zeros((16, 6))
    wout = np.zeros((16, N))
    for f in flag_sets:
        stats[f], wout[f] = ref.dmisft(dat, sig, wmatd, wmatdr, wmatdp, int(f), l1min, d0, fm0)
    w = meshgen.small_block()
    r = Reference(w.mesh)
    el = w.mesh.elanz
    mnr_id = np.arange(1, el + 1, dtype=np.int32)
    mnr_z = (np.arange(el) // 3 + 1).astype(np.int32)
    par_id = ref.bpar(r, w.sigma, el, mnr_id)
    par_z = ref.bpar(r, w.sigma, el // 3 + 1 if el % 3 else el // 3, mnr_z)
    par_step = par_z + 0.05 * (rng.normal(size=len(par_z)) + 1j * rng.normal(size=len(par_z)))
    sig_z = ref.bsigma(r, par_step, el, mnr_z)
    np.savez_compressed(
        OUT / "ref_inversion_steps.npz",
        chk_in_d=np.array([c[0] for c in cin]), chk_in_s=np.array([c[1] for c in cin]),
        chk_in_ld=np.array([c[2] for c in cin], np.int32),
        chk_cdat=np.array([c[0] for c in cout]), chk_csig=np.array([c[1] for c in cout]),
        chk_wdlok=np.array([c[2] for c in cout], np.int32),
        dat=dat, sigmaa=sig, wmatd=wmatd, wmatdr=wmatdr, wmatdp=wmatdp, d0=d0, fm0=fm0,
        l1min=l1min, flags=flag_sets, stats=stats, wmatd_out=wout,
        mnr_z=mnr_z, par_id=par_id, par_z=par_z, par_step=par_step, sigma_z=sig_z)
    print(f"wrote ref_inversion_steps.npz  chkpo2 cases={len(cin)} N={N}")


def blam0_case():
    """blam0 (complex, lfpi and ldc branches) on the small block mesh's bsensi
    Jacobian with the error weights of its voltages, alfx = 1, alfz = 0.5."""
    from oracle import ref
    w = meshgen.small_block()
    r = Reference(w.mesh)
    r.load_data(w.strnr, w.vnr)
    r.rwaven()
    r.precal()
    r.forward(w.sigma, nthreads=8)
    v, _ = r.volt(1)
    J = r.sens(1)
    wt = 1.0 / (0.02 + 1e-4 / np.abs(v)) ** 2
    lam = np.array([ref.blam0(J, wt, m, 1.0, 0.5) for m in (0, 1, 2)])
    np.savez_compressed(OUT / "ref_blam0.npz", w=wt, alfx=1.0, alfz=0.5, lammax=lam)
    print(f"wrote ref_blam0.npz  lammax={lam}")


# data files for rdati: (name, eanz, ldc, error model, text)
RDATI_MODEL = (3.0, 1e-3, 0.5, -0.8, 2.0, 0.1)
RDATI_FILES = [
    ("crt_model", 8, False, "4\n10002 40003 1.2345 -12.5\n20003 50004 0.5 -3.25\n"
                            "30004 60005 2.0e-2 -45.0\n10003 50006 7.5 1.5\n"),
    ("abmn_model", 8, False, "3\n1 2 4 3 0.8 -7.0\n2 3 5 4 1.25d0 -20.5 99 99\n3 4 8 7 3.5E-1 2.0\n"),
    ("crt_indiv", 8, False, "3 T\n10002 40003 1.2345 -12.5 0.05 0.4\n20003 50004 0.5 -3.25 0.01 0.2\n"
                            "30004 60005 2.0e-2 -45.0 0.003 1.5\n1.5 2.0\n"),
    ("dc_indiv", 6, True, "3 .TRUE.\n10002 40003 1.2345 0.05\n20003 50004 0.5 0.01\n"
                          "30004 60005 2.0e-2 0.003\n0.75\n"),
    ("dc_model", 6, True, "3 F\n10002 40003 1.2345 -12.5\n20003 50004 0.5\n30004 60005 2.0e-2 0 0\n"),
    ("spanning", 8, False, "2\n10002 40003\n 1.5 -2.0\n20003\n50004 0.25\n -1.0 extra\n"),
    ("err_dup", 8, False, "1 T\n10002 20003 1.0 -1.0 0.1 0.1\n1.0 1.0\n"),
    ("err_range", 4, False, "1\n10002 40005 1.0 -1.0\n"),
    ("err_phase_error", 8, False, "1 T\n10002 40003 1.0 -1.0 0.1 0.0\n1.0 1.0\n"),
    ("err_resistance", 8, False, "2\n10002 40003 1.0 -1.0\n20003 50004 -0.5 -1.0\n"),
    ("err_phase", 8, False, "1\n10002 40003 1.0 -3200.0\n"),
    ("err_eof", 8, False, "3\n10002 40003 1.0 -1.0\n20003 50004 0.5 -1.0\n"),
]


# synthetic noise models: iseed, nstabw0, nstabm0, nstabpA1, nstabpB, nstabpA2, nstabp0
RDATI_NOISE = [(1, 2.0, 1e-3, 0.3, 0.5, 1.5, 0.2), (12345, 5.0, 0.0, 0.0, 0.0, 0.0, 0.0)]


def rdati_cases():
    """rdati on the data files above (written to tests/golden/rdati/), with
    the error model RDATI_MODEL; error numbers or the arrays rdati fills."""
    from oracle import ref
    d = OUT / "rdati"
    d.mkdir(exist_ok=True)
    out = {}
    for name, eanz, ldc, text in RDATI_FILES:
        f = d / f"{name}.dat"
        f.write_text(text)
        e, arr = ref.rdati(f, eanz, ldc, RDATI_MODEL)
        out[f"{name}__errnr"] = np.int32(e)
        if e == 0:
            for k, v in arr.items():
                out[f"{name}__{k}"] = np.asarray(v)
    for k, (name, eanz, ldc, _) in enumerate([c for c in RDATI_FILES if c[0] in ("crt_model", "abmn_model",
                                                                                  "dc_model")]):
        for j, nz in enumerate(RDATI_NOISE):
            e, arr = ref.rdati(d / f"{name}.dat", eanz, ldc, RDATI_MODEL, noise=nz)
            key = f"{name}__noise{j}"
            out[f"{key}__errnr"] = np.int32(e)
            for kk, v in arr.items():
                out[f"{key}__{kk}"] = np.asarray(v)
    np.savez_compressed(OUT / "ref_rdati.npz", **out)
    print("wrote ref_rdati.npz", {n: int(out[f"{n}__errnr"]) for n, *_ in RDATI_FILES})


def update_case():
    """The reference update (src/update.f90, llam = .false., ltri = 0: cgfac,
    right-hand side J^H W (d - f) - lam R m, cjg -> cjggra, dpar scaled by
    cgfac, par + dpar) on the small block mesh: J and the smoothness matrix
    from the oracle (both pinned to the reference elsewhere), seeded data, m =
    ln sigma, lam = 10, three (ncgmax, eps) settings."""
    from oracle import cpu, ref
    w = meshgen.small_block()
    o = cpu.Oracle(w.mesh)
    o.rwaven(w.strnr, w.vnr)
    o.precal()
    o.forward(w.sigma)
    v, sa = o.volt(w.strnr, w.vnr, 1)
    J = o.sens(1)
    rng = np.random.default_rng(5)
    N, M = J.shape
    dat = sa + 0.01 * (rng.standard_normal(N) + 1j * rng.standard_normal(N))
    wt = 1.0 / (0.02 + 1e-4 / np.abs(v)) ** 2
    sm = cpu.bsmatmreg(o, 40, 20, 1.0, 1.0)
    par = np.log(w.sigma)
    out = dict(dat=dat, w=wt, lam=10.0, nx=40, nz=20)
    for name, ncg, eps in (("a", 10, 0.0), ("b", 20, 1e-4), ("c", 30, 1e-6)):
        dpar, cgfac, cgres, pnew, bd = ref.update(J, wt, dat, sa, par, sm, 10.0, 40, 20, ncg, eps, 1.0)
        out.update({f"{name}_ncg": ncg, f"{name}_eps": eps, f"{name}_dpar": dpar, f"{name}_cgfac": cgfac,
                    f"{name}_cgres": cgres, f"{name}_par": pnew, f"{name}_bdpar": bd})
    np.savez_compressed(OUT / "ref_update.npz", **out)
    print("wrote ref_update.npz")


CASES = {"inversion": inversion_steps, "blam0": blam0_case, "rdati": rdati_cases, "update": update_case,
         "topo_small": topo_small, "triangular": fixture_triangular, "smoothness": smoothness_cases,
         "non_ordered": lambda: fixture("non_ordered", 1)}


def main():
    if not available():
        sys.exit("oracle/_ref/libcrtref.so is not built (make -C oracle ref)")
    if len(sys.argv) > 1:  # named cases only
        for name in sys.argv[1:]:
            CASES[name]()
        return
    synthetic(meshgen.small_block())
    synthetic(meshgen.small_dirichlet())
    synthetic_dc(meshgen.small_block())
    synthetic_dc(meshgen.small_dirichlet())
    synthetic_rb2(meshgen.small_block(), 0)
    synthetic_rb2(meshgen.small_dirichlet(), 1)
    triangular_smoothness()
    fixture("minimal", 1)
    fixture("mixed_surf", 1)
    fixture("decoupling", 0, lsink=True, nsink=373)
    fixture("rect_complex", 1)
    fixture("non_ordered", 1)
    topo_small()
    fixture_triangular()
    smoothness_cases()
    inversion_steps()
    blam0_case()
    rdati_cases()
    update_case()


if __name__ == "__main__":
    main()
