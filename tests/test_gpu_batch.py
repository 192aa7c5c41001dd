"""The graph-captured CG and the uneven batch of Gauss-Newton steps.

* crt_set_graph: the CG iterations replayed from one captured HIP graph give
  the launch-by-launch results bit for bit, on every replay, with eager solves
  on the same handle in between, and a changed parameter recaptures;
* titatest_amd.batch: one Gauss-Newton step on models of different mesh and
  electrode counts, run concurrently (threads, graph-captured CG) against the
  same steps run one by one with launch-by-launch CG (bit for bit), and the
  step against the oracle's chain (bvolti, bsensi, update's right-hand side
  and cgfac, cjggra with a fixed iteration count) to the sensitivity
  tolerance, with the adjoint gradient against J^H W r.  The oracle
  comparison runs 10 CG iterations: with eps = 0 the iterations past
  convergence amplify rounding in the inputs -- on these models the oracle's
  own 20-iteration update moves by up to 6e-10 when J is perturbed by 1e-15
  (3e-15 after 10 iterations) -- so the 20-iteration steps are checked bit for
  bit across the graph, threaded and per-model paths instead.
"""
import numpy as np
import pytest

from conftest import rel_err
from titatest_amd import meshgen

pytestmark = pytest.mark.gpu

STOL = 1e-8


def say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def crtmod(hip_lib):
    from titatest_amd import crt
    return crt


def test_cg_graph_replays_match_eager(crtmod, oracle_lib):
    w = meshgen.small_block()
    say("small_block cg graph vs eager")
    g = crtmod.CRT(w.mesh)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    g.forward(w.sigma)
    v, sa = g.bvolti(w.strnr, w.vnr)
    g.sensitivities(1, to_host=False)
    rng = np.random.default_rng(5)
    N = len(v)
    dat = sa + 0.01 * (rng.standard_normal(N) + 1j * rng.standard_normal(N))
    wgt = 1.0 / (0.02 + 1e-4 / np.abs(v)) ** 2
    sm = g.smoothness(0, 1.0, 1.0, nx=40, nz=20)

    def solve(lam, graph):
        g.set_graph(graph)
        return g.cg_solve(wgt, dat, sa, lam, 0, sm, sm[:, 0], 20, 1e-4, nx=40)

    ref = solve(10.0, False)
    for rep in range(3):
        got = solve(10.0, True)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b)
        eager = solve(10.0, False)  # eager work on the handle between replays
        np.testing.assert_array_equal(eager[0], ref[0])
    ref2 = solve(3.0, False)
    got2 = solve(3.0, True)  # a changed parameter recaptures
    np.testing.assert_array_equal(got2[0], ref2[0])
    assert rel_err(got2[0], ref[0]) > 1e-6


def _small_models(count=2):
    ws = meshgen.uneven_batch(count=8, seed=1006)
    ws = sorted(ws, key=lambda w: w.mesh.elanz)[:count]
    out = []
    for w in ws:
        nx = len(np.unique(w.mesh.sx)) - 1
        out.append((w, nx, w.mesh.elanz // nx))
    return out


def test_uneven_batch_gn_step(crtmod, oracle_lib):
    from titatest_amd import batch
    cases = _small_models(2)
    say("uneven batch", [(w.mesh.sanz, w.mesh.eanz, w.nanz) for w, _, _ in cases])
    mods = [batch.prepare(w, nx, nz, seed=i) for i, (w, nx, nz) in enumerate(cases)]
    seq = [batch.gn_step(m, use_graph=False) for m in mods]
    par, _ = batch.run_batch(mods, threads=2, use_graph=True)
    for (d1, g1, _), (d2, g2, _) in zip(seq, par):
        np.testing.assert_array_equal(d1, d2)
        np.testing.assert_array_equal(g1, g2)
    for (w, nx, nz), m in zip(cases, mods):
        dpar, grad, _ = batch.gn_step(m, ncg=10, use_graph=True)
        o = oracle_lib.Oracle(w.mesh)
        o.rwaven(w.strnr, w.vnr)
        o.precal()
        o.forward(w.sigma)
        v, sa = o.volt(w.strnr, w.vnr, 1)
        J = o.sens(1)
        N = len(v)
        ones = np.ones(N, np.int32)
        gj = J.conj().T @ (m.wgt * (m.dobs - sa))
        assert rel_err(grad, gj) < STOL
        cgfac, bvec = oracle_lib.update_rhs(J, m.wgt, ones, m.dobs, sa, 10.0, 0, m.smatm[:, 0], m.rm)
        dpo, _, ncg = oracle_lib.cjggra(J, m.wgt, ones, cgfac, bvec, 10.0, 0, nx, m.smatm, 10, 0.0)
        assert ncg == 10
        assert rel_err(dpar, dpo * cgfac) < STOL


def test_gn_iterations_reduce_the_misfit(crtmod, oracle_lib):
    """Three fixed-lambda Gauss-Newton iterations (forward, misfit, J, CG,
    bpar -> update -> bsigma, all through the C-ABI) move the model towards
    the data: the misfit falls, and each iteration's misfit is the oracle's
    dmisft of the same modelled data."""
    from titatest_amd import batch
    from oracle import cpu
    (w, nx, nz), = _small_models(1)
    say("gn iterations", w.mesh.sanz, w.mesh.eanz, w.nanz)
    m = batch.prepare(w, nx, nz, seed=3, contrast=2.0)
    # from the homogeneous model of the same mean log conductivity; lam keeps
    # the steps in the linear range (no line search); the oracle's chain on
    # this model gives nrmsd 43.5, 13.3, 10.5, 9.48
    s0 = np.full_like(w.sigma, np.exp(np.mean(np.log(w.sigma))))
    hist, sig = batch.gn_iterate(m, iters=3, lam=1e4, sigma0=s0)
    nr = [h[0] for h in hist]
    assert nr[0] > nr[1] > nr[2] > nr[3], nr
    assert abs(nr[0] - 43.4673) < 1e-3 and nr[3] < 0.25 * nr[0], nr
    assert np.all(np.isfinite(sig)) and not np.array_equal(sig, w.sigma)
    g = m.g
    g.forward(sig)
    _, sa = g.bvolti(w.strnr, w.vnr)
    so, _, _ = cpu.dmisft(m.dobs, sa, m.wgt, m.wgt, m.wgt)
    assert abs(nr[3] - so[0]) <= 1e-13 * so[0]
    assert abs(hist[3][1] - so[1]) <= 1e-13 * so[1] and abs(hist[3][2] - so[2]) <= 1e-13 * so[2]


def test_forward_many_equals_per_model_forwards(crtmod):
    """crt_forward_many over models of different bands and mesh types -- four
    uneven quadrilateral models, a terrain-following triangular mesh with
    lsytop, a band of 132 (the wide form: window in global memory) and a
    Dirichlet model (one factor per wavenumber, its own forward) -- gives
    each model the voltages (both modes, 1e-12) and Jacobian (1e-10) of its
    own crt_forward: the persistent launches run the same per-system algorithm in
    the same operation order, compiled separately (the compiler's choice of
    fused multiply-adds may differ by a rounding)."""
    from titatest_amd import batch
    ws = sorted(meshgen.uneven_batch(count=8, seed=1006), key=lambda w: w.mesh.elanz)[:4]
    tri = meshgen.topo_triangles(nx_nodes=41, nz_nodes=20, eanz=16)
    ws += [tri, meshgen.large_grf(nx_side=2, n_uniform=12, nz=130, eanz=12),
           meshgen.large_grf(nx_side=4, n_uniform=20, nz=12, eanz=16, side_type=13)]
    lsy = [w is tri for w in ws]
    say("forward_many over", [(w.mesh.sanz, w.mesh.mb) for w in ws])
    gm = [crtmod.CRT(w.mesh, lsytop=t) for w, t in zip(ws, lsy)]
    gr = [crtmod.CRT(w.mesh, lsytop=t) for w, t in zip(ws, lsy)]
    for g, w in zip(gm + gr, ws + ws):
        g.rwaven(w.strnr, w.vnr)
        g.precal()
    batch.forward_many(gm, [w.sigma for w in ws])
    for g1, g2, w in zip(gm, gr, ws):
        g2.forward(w.sigma)
        for mode in (0, 1):
            v1, s1 = g1.voltages(w.strnr, w.vnr, mode)
            v2, s2 = g2.voltages(w.strnr, w.vnr, mode)
            assert rel_err(v1, v2) < 1e-12 and rel_err(s1, s2) < 1e-12
        # J forms differences of potentials: rounding shows at ~1e-12
        assert rel_err(g1.sensitivities(1), g2.sensitivities(1)) < 1e-10


def test_persistent_batch_equals_per_model_steps(crtmod):
    """The persistent batch's Gauss-Newton steps against each model's own
    step: gradients to 1e-10, the 10-iteration CG steps to the sensitivity
    tolerance (the forwards agree to rounding; see above)."""
    from titatest_amd import batch
    cases = _small_models(3)
    say("persistent batch", [(w.mesh.sanz, w.mesh.eanz) for w, _, _ in cases])
    mods = [batch.prepare(w, nx, nz, seed=i) for i, (w, nx, nz) in enumerate(cases)]
    seq = [batch.gn_step(m, ncg=10, use_graph=False) for m in mods]
    per, _ = batch.run_batch_persistent(mods, threads=3, ncg=10, use_graph=True)
    for (d1, g1, c1), (d2, g2, c2) in zip(seq, per):
        assert rel_err(g1, g2) < 1e-10
        assert rel_err(d1, d2) < STOL
        assert int(c1[0]) == int(c2[0])


def test_autotuner_choice_and_cache(crtmod, tmp_path):
    """The autotuner times the four launch choices on the model, applies the
    fastest, records it by model shape, and a second handle of that shape
    takes it from the cache; every choice gives the default voltages to
    rounding."""
    from titatest_amd import autotune
    w = meshgen.small_block()
    g = crtmod.CRT(w.mesh)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    g.forward(w.sigma)
    v0, _ = g.bvolt(w.strnr, w.vnr)
    cache = tmp_path / "tune.json"
    choice, times = autotune.tune(g, w.sigma, cache=cache, reps=1)
    say("autotune choice", choice, times)
    assert set(times) == {(t, d) for t in (256, 512) for d in (8, 16)}
    assert choice == min(times, key=times.get)
    for t, d in times:
        g.set_tuning(t, d)
        g.forward(w.sigma)
        v, _ = g.bvolt(w.strnr, w.vnr)
        assert rel_err(v, v0) < 1e-12
    g2 = crtmod.CRT(w.mesh)
    g2.rwaven(w.strnr, w.vnr)
    g2.precal()
    choice2, times2 = autotune.tune(g2, w.sigma, cache=cache)
    assert choice2 == choice and times2 == {}
    with pytest.raises(crtmod.CRTError):
        g2.set_tuning(384, 8)
