"""Band widths above 124: the factorisation with its trailing window in a
global scratch (panels and staged rows in LDS), and the substitutions with
four register slot classes per lane (a 256-row window).

* the reference's own triangular modelling case (tests/modelling_triangular:
  band width 199, 1,489 nodes, 215 electrodes) against the golden vectors
  written from the reference's routines compiled from source;
* the widest band taken, mb = 248, on a terrain-following triangular mesh
  with lsytop, against the oracle and the long-double solution; mb = 249
  refused with 1003;
* the shared-factor forward (no mixed boundaries) at mb = 142.

Tolerances: voltages 1e-10 relative in magnitude, 1e-9 rad in phase;
potentials 1e-9 (the wide systems' own conditioning, measured against the
long-double solution); sensitivities 1e-8.
"""
import numpy as np
import pytest

from conftest import FIXTURES, golden, rel_err
from titatest_amd import formats as F
from titatest_amd import meshgen

pytestmark = pytest.mark.gpu

VTOL = 1e-10
PTOL = 1e-9
STOL = 1e-8


def say(*a):
    print("[gpu-case]", *a, flush=True)


def check_volt(v, vref):
    mag = np.abs(np.abs(v) - np.abs(vref)) / np.abs(vref)
    dph = np.abs(np.angle(v / vref))
    assert mag.max() < VTOL, mag.max()
    assert dph.max() < PTOL, dph.max()


@pytest.fixture(scope="module")
def crtmod(hip_lib):
    from titatest_amd import crt
    return crt


def test_reference_triangular_case(crtmod):
    d = FIXTURES / "modelling_triangular"
    m = F.read_elem(d / "elem.dat")
    m.enr = F.read_elec(d / "elec.dat")
    sig = F.read_rho(d / "rho.dat", m.elanz)
    st, vn = F.read_config(d / "config.dat", m.eanz)
    ge = golden("modelling_triangular")
    say("reference modelling_triangular mb=", m.mb, "n=", m.sanz, "E=", m.eanz)
    assert m.mb == 199
    g = crtmod.CRT(m)
    g.rwaven(st, vn)
    np.testing.assert_array_equal(g.kwn, ge["kwn"])
    g.precal()
    kpot, hpot = g.forward(sig, want_kpot=True, want_hpot=True)
    for (l, k), ref in zip(ge["systems"], ge["kpot_sys"]):
        assert rel_err(kpot[:, l, k], ref) < PTOL, (l, k)
    enr = m.enr - 1
    assert rel_err(hpot[enr], ge["hpot_elec"]) < PTOL
    v, sa = g.bvolt(st, vn)
    check_volt(v, ge["volt"])
    v1, sa1 = g.bvolti(st, vn)
    assert np.max(np.abs(sa1 - ge["sigmaa1"])) < VTOL
    s1 = g.bsensi()
    assert rel_err(s1[:4], ge["sens1"]) < STOL


def test_widest_band_248_triangles_with_topography(crtmod, oracle_lib):
    w = meshgen.topo_triangles(nx_nodes=24, nz_nodes=247, eanz=16, length=30.0, depth=40.0, seed=1031)
    say("widest band mb=", w.mesh.mb, "n=", w.mesh.sanz)
    assert w.mesh.mb == 248
    g = crtmod.CRT(w.mesh, lsytop=True)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    (kpot,) = g.forward(w.sigma, want_kpot=True)
    o = oracle_lib.Oracle(w.mesh, lsytop=True)
    o.rwaven(w.strnr, w.vnr)
    o.precal()
    E = w.mesh.eanz
    for s in (0, 4 * E + 7, (g.K - 1) * E + E - 1):
        o.forward(w.sigma, sys_range=(s, s + 1))
        k, l = divmod(s, E)
        assert rel_err(kpot[:, l, k], o.kpot[:, l, k]) < PTOL, (k, l)
    kl = o.solve_ld(w.sigma, 0, 0)
    e_hip, e_orc = rel_err(kpot[:, 0, 0], kl), rel_err(o.kpot[:, 0, 0], kl)
    say("k=0 l=0 distance from the long-double solution: hip", e_hip, "oracle", e_orc)
    assert e_hip < max(3 * e_orc, 1e-13), (e_hip, e_orc)
    del g
    w2 = meshgen.topo_triangles(nx_nodes=24, nz_nodes=248, eanz=16, length=30.0, depth=40.0, seed=1031)
    assert w2.mesh.mb == 249
    g2 = crtmod.CRT(w2.mesh, lsytop=True)
    g2.rwaven(w2.strnr, w2.vnr)
    g2.precal()
    with pytest.raises(crtmod.CRTError) as e:
        g2.forward(w2.sigma)
    assert e.value.code == 1003


def test_wide_band_shared_factor(crtmod, oracle_lib):
    """No mixed boundaries at mb = 142: one factor per wavenumber in the
    wide-band form, the column-band copy, and every source's forward sweep and
    back substitution with four slot classes; all potentials and the
    voltages against the oracle."""
    w = meshgen.large_grf(nx_side=4, n_uniform=24, nz=140, eanz=24, side_type=13)
    say("dirichlet wide band mb=", w.mesh.mb, "n=", w.mesh.sanz)
    assert w.mesh.mb == 142
    g = crtmod.CRT(w.mesh)
    g.rwaven(w.strnr, w.vnr)
    assert g.precal() == (0, 1)
    (kpot,) = g.forward(w.sigma, want_kpot=True)
    o = oracle_lib.Oracle(w.mesh)
    o.rwaven(w.strnr, w.vnr)
    o.precal()
    o.forward(w.sigma)
    assert rel_err(kpot, o.kpot) < PTOL
    v, _ = g.bvolt(w.strnr, w.vnr)
    vo, _ = o.volt(w.strnr, w.vnr, 0)
    check_volt(v, vo)
