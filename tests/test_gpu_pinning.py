"""The HIP path on terrain and triangles, and the smoothness builders, against
the reference's golden vectors (written from the reference's routines
compiled from source):

* a terrain-following triangular mesh with lsytop (elem3 on sheared
  triangles, beta's image sources mirrored about bsytop's height): element
  matrices and beta on the device, potentials, voltages both modes,
  sensitivities both modes;
* bsmatmreg on a graded mesh and bsmatmtri with decoupled edges (the
  reference's own decouplings.dat files, and seeded pairs).

Tolerances: potentials and voltages 1e-10 relative in magnitude, 1e-9 rad in
phase; sensitivities 1e-8 relative.
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


@pytest.fixture(scope="module")
def topo(crtmod):
    w = meshgen.topo_triangles(nx_nodes=31, nz_nodes=12, eanz=16, length=30.0, depth=12.0, seed=1013)
    say("topo_tri_small lsytop n=", w.mesh.sanz, "mb=", w.mesh.mb)
    g = crtmod.CRT(w.mesh, lsytop=True)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    return w, g


def test_topo_precal_device(topo):
    w, g = topo
    ge = golden("topo_tri_small")
    np.testing.assert_array_equal(g.kwn, ge["kwn"])
    elbg, relbg, kg = g.get_precal()
    m = w.mesh
    nimn = m.maxnodes * (m.maxnodes + 1) // 2
    assert rel_err(elbg.reshape((m.elanz, nimn, g.K), order="F"), ge["elbg"]) < 1e-13
    btyp = np.concatenate([np.full(n, t) for t, n in zip(m.typ, m.nelanz) if t > 10])
    mixed = btyp == 11
    kgr = kg.reshape((m.relanz, m.eanz, g.K), order="F")[mixed][:64]
    assert rel_err(kgr, ge["kg"]) < 1e-12


def test_topo_forward_volt_sens(topo):
    w, g = topo
    ge = golden("topo_tri_small")
    say("topo_tri_small forward + volt + sens")
    kpot, hpot = g.forward(w.sigma, want_kpot=True, want_hpot=True)
    enr = w.mesh.enr - 1
    assert rel_err(kpot[enr], ge["kpot_elec"]) < VTOL
    assert rel_err(kpot[:, 0, :], ge["kpot_col"]) < VTOL
    assert rel_err(hpot[enr], ge["hpot_elec"]) < VTOL
    v, sa = g.bvolt(w.strnr, w.vnr)
    check_volt(v, ge["volt"])
    s0 = g.bsens()
    assert rel_err(s0[:8], ge["sens0"]) < STOL
    v1, sa1 = g.bvolti(w.strnr, w.vnr)
    assert np.max(np.abs(sa1 - ge["sigmaa1"])) < VTOL
    s1 = g.bsensi()
    assert rel_err(s1[:8], ge["sens1"]) < STOL


def test_bsmatmreg_graded_mesh(crtmod):
    ge = golden("smoothness")
    w = meshgen.large_grf(nx_side=5, n_uniform=14, nz=9, eanz=12, seed=1021)
    nx, nz = int(ge["reg_nx"]), int(ge["reg_nz"])
    say("bsmatmreg graded", nx, "x", nz)
    g = crtmod.CRT(w.mesh)
    # host arithmetic may contract products into fused multiply-adds
    assert rel_err(g.smoothness(0, 1.0, 0.5, nx=nx, nz=nz), ge["reg_smatm"]) < 1e-14
    assert rel_err(g.smoothness(0, 3.0, 0.25, nx=nx, nz=nz), ge["reg_smatm_b"]) < 1e-14


def test_decoupled_triangular_smoothness(crtmod):
    ge = golden("smoothness")
    d = FIXTURES / "decoupling"
    m = F.read_elem(d / "elem.dat")
    m.enr = F.read_elec(d / "elec.dat")
    say("bsmatmtri with decouplings.dat")
    g = crtmod.CRT(m)
    for tag in ("line", "full"):
        ed, es = F.read_decouplings(d / f"decouplings_{tag}.dat")
        g.set_decouplings(ed, es)
        sm, nb = g.smoothness(1, 1.0, 1.0)
        np.testing.assert_array_equal(nb, ge["dec_nachbar"])
        assert rel_err(sm, ge[f"dec_{tag}_smatm"]) < 1e-14
    g.set_decouplings(np.zeros((0, 2), np.int32), np.zeros(0))
    sm0, _ = g.smoothness(1, 1.0, 1.0)
    assert rel_err(sm0, ge["dec_line_smatm"]) > 1e-3
    with pytest.raises(crtmod.CRTError):
        g.set_decouplings(np.array([[1, m.elanz + 1]]), np.array([0.5]))
    w = meshgen.topo_triangles(nx_nodes=25, nz_nodes=10, eanz=12)
    g2 = crtmod.CRT(w.mesh)
    g2.set_decouplings(ge["tri_edecoup"], ge["tri_edecstr"])
    sm, _ = g2.smoothness(1, 1.0, 0.5)
    assert rel_err(sm, ge["tri_smatm"]) < 1e-14
