"""Gram form of the explicit Jacobian (crt_set_sens_form 2): per element, the
E x E matrix of every source pair's bilinear term, summed over the
wavenumbers on the FP64 matrix cores, then one double difference per
measurement.  Against the oracle's bsens / bsensi / bsendc and the
reference's golden rows at 1e-8 relative, and against the pair form
(crt_set_sens_form 1) to rounding:

* the 861-node block (20 electrodes, one tile row), both modes, golden rows;
* pole configurations (electrode 0 in a current or potential pair);
* random quadrupoles on 37 electrodes (a partly filled third tile row, three
  tiles per wavefront, Gauss's three products), 96 (all 21 tiles, six per
  wavefront, four products) and 128 (36 tiles, nine per wavefront), both
  modes;
* no wavenumber sum (swrtr 0, one wavenumber);
* the DC twin (bsendc);
* the 240 x 80 workload (96 electrodes, 3,090 quadrupoles, 18 of the 21
  tiles formed, five per wavefront), where the built-in rule takes the Gram
  form: equal to the pair form.
"""
import numpy as np
import pytest

from conftest import golden, rel_err
from titatest_amd import meshgen

pytestmark = pytest.mark.gpu

STOL = 1e-8
AGREE = 1e-12  # the two forms: same sums, other rounding


def say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def crtmod(hip_lib):
    from titatest_amd import crt
    return crt


def _setup(crt, oracle_lib, mesh, st, vn, sigma):
    g = crt.CRT(mesh)
    g.rwaven(st, vn)
    g.precal()
    g.forward(sigma)
    o = oracle_lib.Oracle(mesh)
    o.rwaven(st, vn)
    o.precal()
    o.forward(sigma)
    return g, o


def _both_forms(g, mode):
    g.set_sens_form(1)
    jp = g.sensitivities(mode)
    g.set_sens_form(2)
    jg = g.sensitivities(mode)
    g.set_sens_form(0)
    return jp, jg


def test_gram_small_block_both_modes(crtmod, oracle_lib):
    w = meshgen.small_block()
    say("gram small_block E=20")
    g, o = _setup(crtmod, oracle_lib, w.mesh, w.strnr, w.vnr, w.sigma)
    ge = golden("small_block")
    for mode in (0, 1):
        g.voltages(w.strnr, w.vnr, mode)
        o.volt(w.strnr, w.vnr, mode)
        jp, jg = _both_forms(g, mode)
        assert rel_err(jg[:8], ge[f"sens{mode}"]) < STOL
        assert rel_err(jg, o.sens(mode)) < STOL
        assert rel_err(jg, jp) < AGREE


def test_gram_pole_configurations(crtmod, oracle_lib):
    w = meshgen.small_block()
    st = np.array([10000, 20000, 30004, 40005], np.int32)
    vn = np.array([50000, 60007, 90000, 80000], np.int32)
    say("gram pole configs")
    g, o = _setup(crtmod, oracle_lib, w.mesh, st, vn, w.sigma)
    for mode in (0, 1):
        g.voltages(st, vn, mode)
        o.volt(st, vn, mode)
        jp, jg = _both_forms(g, mode)
        assert rel_err(jg, o.sens(mode)) < STOL
        assert rel_err(jg, jp) < AGREE


@pytest.mark.parametrize("E,nmeas", [(37, 700), (96, 1500), (128, 1500)])
def test_gram_random_quadrupoles(crtmod, oracle_lib, E, nmeas):
    w = meshgen.large_grf(nx_side=4, n_uniform=E, nz=8, eanz=E)
    rng = np.random.default_rng(E)
    q = np.array([rng.choice(E, 4, replace=False) + 1 for _ in range(nmeas)])
    st, vn = q[:, 0] * 10000 + q[:, 1], q[:, 2] * 10000 + q[:, 3]
    say("gram random quadrupoles E=", E, "measurements", nmeas, "n=", w.mesh.sanz)
    g, o = _setup(crtmod, oracle_lib, w.mesh, st, vn, w.sigma)
    for mode in (0, 1):
        g.voltages(st, vn, mode)
        o.volt(st, vn, mode)
        jp, jg = _both_forms(g, mode)
        assert rel_err(jg, o.sens(mode)) < STOL
        assert rel_err(jg, jp) < AGREE


def test_gram_without_wavenumber_sum(crtmod, oracle_lib):
    w = meshgen.small_block()
    say("gram swrtr 0")
    g = crtmod.CRT(w.mesh)
    # one wavenumber, its potentials taken as they are (no back-transform);
    # 0.05 rather than 0 keeps the mixed boundary's Bessel terms finite
    g.set_wavenumbers([0.05], [1.0], swrtr=0)
    g.precal()
    g.forward(w.sigma)
    o = oracle_lib.Oracle(w.mesh)
    o.set_waven([0.05], [1.0], swrtr=0)
    o.precal()
    o.forward(w.sigma)
    for mode in (0, 1):
        g.voltages(w.strnr, w.vnr, mode)
        o.volt(w.strnr, w.vnr, mode)
        jp, jg = _both_forms(g, mode)
        assert rel_err(jg, o.sens(mode)) < STOL
        assert rel_err(jg, jp) < AGREE


def test_gram_dc_twin(crtmod, oracle_lib):
    w = meshgen.small_block()
    say("gram bsendc small_block")
    g = crtmod.CRT(w.mesh)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    o = oracle_lib.Oracle(w.mesh)
    o.rwaven(w.strnr, w.vnr)
    o.precal()
    g.forward_dc(w.sigma)
    o.forward_dc(w.sigma)
    g.bvolti(w.strnr, w.vnr)
    o.volt(w.strnr, w.vnr, 1)
    g.set_sens_form(1)
    sp = g.bsendc()
    g.set_sens_form(2)
    sg = g.bsendc()
    so = o.sens_dc()
    assert rel_err(sg, so) < STOL
    assert rel_err(sg[:8], golden("small_block_dc")["sensdc"]) < STOL
    assert rel_err(sg, sp) < AGREE


def test_gram_headline_workload_equals_pair_form(crtmod):
    w = meshgen.large_grf()
    say("gram 240x80 E=96 N=3090: built-in rule against the pair form")
    g = crtmod.CRT(w.mesh)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    g.forward(w.sigma)
    for mode in (0, 1):
        g.voltages(w.strnr, w.vnr, mode)
        g.set_sens_form(1)
        jp = g.sensitivities(mode)
        g.set_sens_form(0)
        j0 = g.sensitivities(mode)
        g.set_sens_form(2)
        jg = g.sensitivities(mode)
        g.set_sens_form(0)
        np.testing.assert_array_equal(j0, jg)
        assert rel_err(jg, jp) < AGREE


def test_gram_form_refused_above_128_electrodes(crtmod):
    w = meshgen.large_grf(nx_side=4, n_uniform=130, nz=6, eanz=130)
    g = crtmod.CRT(w.mesh)
    with pytest.raises(crtmod.CRTError) as ex:
        g.set_sens_form(2)
    assert ex.value.code == 1000
    g.set_sens_form(1)
    g.set_sens_form(0)
