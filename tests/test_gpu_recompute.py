"""Per-wavenumber recomputation of the potentials (crt_set_recompute): the
forward keeps only each pass's contribution to the electrode rows of the
back-transform, the adjoint gradient re-solves each pass's systems.

* the voltages equal the stored mode's bit for bit (the same sums in the
  same order), for passes of 1, 3 (an uneven last pass) and all wavenumbers,
  on a mesh with mixed boundaries (one factor per system) and one without
  (one factor per wavenumber serving every source);
* the adjoint gradient equals the stored mode's to 1e-12 relative in both
  modes (the element sums are added pass by pass instead of in one run);
* the entry points that need every wavenumber's potentials refuse with 1002
  (1000 for the forward's kpot and hpot outputs);
* the headline 240x80 workload in passes of 6 wavenumbers: voltages bit for
  bit, gradient to 1e-12;
* the DC twin in passes: its voltages equal the stored DC forward's bit for
  bit, and bsendc (which needs every potential) refuses with 1002.
"""
import numpy as np
import pytest

from conftest import rel_err
from titatest_amd import meshgen

pytestmark = pytest.mark.gpu


def say(*a):
    print("[gpu-case]", *a, flush=True)


def _run(crt, w, passes, seed=5):
    g = crt.CRT(w.mesh)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    if passes:
        g.set_recompute(passes)
    g.forward(w.sigma)
    v, _ = g.bvolti(w.strnr, w.vnr)
    rng = np.random.default_rng(seed)
    wt = rng.uniform(0.5, 2.0, len(w.strnr))
    r = rng.normal(size=len(w.strnr)) + 1j * rng.normal(size=len(w.strnr))
    g1 = g.adjoint_grad(wt, r, mode=1)
    g0 = g.adjoint_grad(wt, r, mode=0)
    return g, v, g1, g0


@pytest.mark.parametrize("mesh", ["small_block", "small_dirichlet"])
def test_recompute_matches_stored(hip_lib, mesh):
    from titatest_amd import crt
    w = getattr(meshgen, mesh)()
    say("recompute", mesh, "stored")
    _, v, g1, g0 = _run(crt, w, 0)
    for passes in (1, 3, 64):
        say("recompute", mesh, "passes", passes)
        gr, vr, g1r, g0r = _run(crt, w, passes)
        np.testing.assert_array_equal(vr, v)
        assert rel_err(g1r, g1) < 1e-12
        assert rel_err(g0r, g0) < 1e-12
    # what needs every wavenumber's potentials
    assert gr.lib.crt_sensitivities(gr.h, 1, None) == 1002
    with pytest.raises(crt.CRTError) as e:
        gr.forward(w.sigma, want_kpot=True)
    assert e.value.code == 1000


def test_recompute_headline_workload(hip_lib):
    import time
    from titatest_amd import crt
    w = meshgen.large_grf()
    say("recompute 240x80 stored")
    t0 = time.perf_counter()
    _, v, g1, g0 = _run(crt, w, 0)
    t1 = time.perf_counter()
    say("recompute 240x80 passes of 6")
    _, vr, g1r, g0r = _run(crt, w, 6)
    t2 = time.perf_counter()
    say(f"stored {t1 - t0:.2f} s, recomputing {t2 - t1:.2f} s (setup included)")
    np.testing.assert_array_equal(vr, v)
    assert rel_err(g1r, g1) < 1e-12
    assert rel_err(g0r, g0) < 1e-12


@pytest.mark.parametrize("mesh", ["small_block", "small_dirichlet"])
def test_recompute_dc_twin(hip_lib, mesh):
    from titatest_amd import crt
    w = getattr(meshgen, mesh)()
    say("recompute DC", mesh)
    g = crt.CRT(w.mesh)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    g.forward_dc(w.sigma)
    v, _ = g.bvolti(w.strnr, w.vnr)
    for passes in (1, 4):
        gr = crt.CRT(w.mesh)
        gr.rwaven(w.strnr, w.vnr)
        gr.precal()
        gr.set_recompute(passes)
        gr.forward_dc(w.sigma)
        vr, _ = gr.bvolti(w.strnr, w.vnr)
        np.testing.assert_array_equal(vr, v)
        out = np.zeros(gr.nanz * gr.elanz)
        assert gr.lib.crt_sensitivities_dc(gr.h, crt._ptr(out)) == 1002
        with pytest.raises(crt.CRTError) as e:
            gr.forward_dc(w.sigma, want_kpot=True)
        assert e.value.code == 1000
