"""Full-size checks of the remaining workloads, each one a wrong kernel fails.

* Terrain-following triangular mesh (257 x 100 nodes = 25,700, 50,688
  triangles, mb 101, 128 electrodes, 7,074 dipole-dipole quadrupoles, 15
  wavenumbers, lsytop): every (k, l) system of the sources that three
  measurements spread over the line need (their current and potential
  electrodes, all wavenumbers) against the oracle; those measurements'
  voltages, both modes; their rows of the explicit Jacobian, both modes;
  the adjoint gradient against J^H (w o r) from the full 7,074 x 50,688
  Jacobian, both modes; five fixed-lambda Gauss-Newton iterations with the
  triangular smoothness operator and the explicit-J CG, the misfit falling
  every iteration and the last one equal to the oracle's dmisft of the same
  modelled data.
* The 64-model uneven batch: graph-captured, threaded steps equal to the
  launch-by-launch steps run one model after another, bit for bit; on the
  two largest models J, the adjoint gradient and a 10-iteration step against
  the oracle's chain, and a step in CRTomo's own CG mode (eps = 1e-4, at most
  20 iterations) against the oracle's CG on the same J: the first ten
  residuals to 1e-8, the step within 30 times the spread of the oracle's own
  step under 1e-15 perturbations of J (on these models 1e-6 and 2e-4: an
  unconverged 20th iterate is not determined better in FP64).
* The 240 x 80 model with Dirichlet sides (where the reference's sensitivity
  is the exact derivative of the discrete forward): the adjoint gradient of
  the weighted misfit against central differences of the forward, step 1e-6
  in ln sigma, real and imaginary directions, on 64 seeded cells.

Tolerances: potentials 1e-9 (their conditioning), voltages 1e-10 relative in
magnitude and 1e-9 rad in phase, sensitivities and gradients 1e-8; central
differences 1e-6 of the gradient's largest entry (their truncation and
round-off).
"""
import numpy as np
import pytest

from conftest import rel_err
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


def _electrodes(st, vn):
    """(A, B, M, N) of each measurement, 1-based, 0 = absent."""
    return np.stack([st // 10000, st % 10000, vn // 10000, vn % 10000], 1)


def test_triangular_terrain_full_size(crtmod, oracle_lib):
    w = meshgen.topo_triangles()
    m = w.mesh
    say("terrain triangles n=", m.sanz, "mb=", m.mb, "E=", m.eanz, "N=", w.nanz)
    g = crtmod.CRT(m, lsytop=True)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    (kpot,) = g.forward(w.sigma, want_kpot=True)
    N = w.nanz
    sel = np.array([0, N // 2, N - 1])
    st, vn = w.strnr[sel], w.vnr[sel]
    srcs = sorted({int(e) - 1 for e in _electrodes(st, vn).ravel() if e > 0})
    o = oracle_lib.Oracle(m, lsytop=True)
    o.rwaven(w.strnr, w.vnr)
    np.testing.assert_array_equal(g.kwn, o.kwn)
    o.precal()
    say("oracle systems:", len(srcs) * o.K)
    o.forward_systems(w.sigma, [(l, k) for l in srcs for k in range(o.K)])
    for l in srcs:
        for k in range(o.K):
            assert rel_err(kpot[:, l, k], o.kpot[:, l, k]) < PTOL, (l, k)
    # the back-transform of the computed sources (rtrafo, src/rtrafo.f90:77-87)
    hp = np.zeros((m.sanz, m.eanz), np.complex128)
    for l in srcs:
        acc = np.zeros(m.sanz, np.complex128)
        for k in range(o.K):
            acc = acc + o.kwnwi[k] * o.kpot[:, l, k]
        hp[:, l] = acc / np.pi
    o.hpot = hp
    del kpot
    v_all, _ = g.bvolt(w.strnr, w.vnr)
    vo, _ = o.volt(st, vn, 0)
    check_volt(v_all[sel], vo)
    rng = np.random.default_rng(1042)
    wgt = rng.uniform(0.5, 2.0, N)
    r = rng.standard_normal(N) + 1j * rng.standard_normal(N)
    for mode in (0, 1):
        g.voltages(w.strnr, w.vnr, mode)
        J = g.sensitivities(mode)
        o.volt(st, vn, mode)
        assert rel_err(J[sel], o.sens(mode)) < STOL, mode
        ga = g.adjoint_grad(wgt, r, mode=mode)
        assert rel_err(ga, J.conj().T @ (wgt * r)) < STOL, mode
        del J


def test_triangular_terrain_gauss_newton(crtmod):
    """Five fixed-lambda iterations (forward, dmisft, explicit J, CG with the
    triangular operator, bpar -> update -> bsigma) from the homogeneous model
    of the same mean log conductivity; lam = 1e4 keeps the full steps in the
    linear range (no line search)."""
    from oracle import cpu
    from titatest_amd import batch
    w = meshgen.topo_triangles()
    say("terrain triangles Gauss-Newton, 5 iterations")
    mod = batch.prepare_tri(w, seed=1041, contrast=2.0)
    s0 = np.full_like(w.sigma, np.exp(np.mean(np.log(w.sigma))))
    hist, sig = batch.gn_iterate(mod, iters=5, lam=1e4, sigma0=s0)
    nr = [h[0] for h in hist]
    say("nrmsd", nr)
    assert all(a > b for a, b in zip(nr, nr[1:])), nr
    assert nr[-1] < 0.25 * nr[0], nr
    g = mod.g
    g.forward(sig)
    _, sa = g.bvolti(w.strnr, w.vnr)
    so, _, _ = cpu.dmisft(mod.dobs, sa, mod.wgt, mod.wgt, mod.wgt)
    for i in range(3):
        assert abs(hist[-1][i] - so[i]) <= 1e-13 * abs(so[i]), (i, hist[-1][i], so[i])


def test_uneven_batch_of_64(crtmod, oracle_lib):
    from titatest_amd import batch
    ws = meshgen.uneven_batch(count=64, seed=1006)
    mods = []
    for i, w in enumerate(ws):
        nx = len(np.unique(w.mesh.sx)) - 1
        mods.append(batch.prepare(w, nx, w.mesh.elanz // nx, seed=1003 + i))
    say("64-model batch: cells", sum(w.mesh.elanz for w in ws))
    seq = [batch.gn_step(m, use_graph=False) for m in mods]
    par, _ = batch.run_batch(mods, threads=8, use_graph=True)
    for (d1, g1, c1), (d2, g2, c2) in zip(seq, par):
        np.testing.assert_array_equal(d1, d2)
        np.testing.assert_array_equal(g1, g2)
        np.testing.assert_array_equal(c1, c2)
    for m in sorted(mods, key=lambda q: q.w.mesh.elanz)[-2:]:
        w = m.w
        say("oracle chain on model n=", w.mesh.sanz, "E=", w.mesh.eanz, "N=", w.nanz)
        g = m.g
        g.forward(w.sigma)
        _, sa = g.bvolti(w.strnr, w.vnr)
        grad = g.adjoint_grad(m.wgt, m.dobs - sa, mode=1)
        Jd = g.sensitivities(1)
        o = oracle_lib.Oracle(w.mesh)
        o.rwaven(w.strnr, w.vnr)
        o.precal()
        o.forward(w.sigma)
        v, sao = o.volt(w.strnr, w.vnr, 1)
        J = o.sens(1)
        ones = np.ones(len(v), np.int32)
        assert rel_err(Jd, J) < STOL
        assert rel_err(grad, J.conj().T @ (m.wgt * (m.dobs - sao))) < STOL
        # 10 iterations against the oracle's whole chain
        dpar, cgfac, _ = g.cg_solve(m.wgt, m.dobs, sa, 10.0, 0, m.smatm, m.smatm[:, 0], 10, eps=0.0,
                                    nx=m.nx, rm=m.rm)
        cgf, bvec = oracle_lib.update_rhs(J, m.wgt, ones, m.dobs, sao, 10.0, 0, m.smatm[:, 0], m.rm)
        dpo, _, _ = oracle_lib.cjggra(J, m.wgt, ones, cgf, bvec, 10.0, 0, m.nx, m.smatm, 10, 0.0)
        assert rel_err(cgfac, cgf) < 1e-12
        assert rel_err(dpar, dpo * cgf) < STOL
        # CRTomo's CG mode (eps = 1e-4, at most 20 iterations) on the device's
        # J and apparent values.  Before convergence the later iterations
        # amplify rounding in their inputs, so the step is held to the spread
        # the oracle's own CG shows when J is perturbed by 1e-15 (relative),
        # and the residual histories to each other.
        dpar, cgfac, cgres = g.cg_solve(m.wgt, m.dobs, sa, 10.0, 0, m.smatm, m.smatm[:, 0], 20, eps=1e-4,
                                        nx=m.nx, rm=m.rm)
        cgf, bvec = oracle_lib.update_rhs(Jd, m.wgt, ones, m.dobs, sa, 10.0, 0, m.smatm[:, 0], m.rm)
        dpo, cgro, ncg = oracle_lib.cjggra(Jd, m.wgt, ones, cgf, bvec, 10.0, 0, m.nx, m.smatm, 20, 1e-4)
        rng = np.random.default_rng(1043)
        spread = 0.0
        for _ in range(3):
            xi = rng.standard_normal(Jd.shape) + 1j * rng.standard_normal(Jd.shape)
            Jp = Jd * (1.0 + 1e-15 * xi)
            cgp, bvp = oracle_lib.update_rhs(Jp, m.wgt, ones, m.dobs, sa, 10.0, 0, m.smatm[:, 0], m.rm)
            dpp, _, _ = oracle_lib.cjggra(Jp, m.wgt, ones, cgp, bvp, 10.0, 0, m.nx, m.smatm, 20, 1e-4)
            spread = max(spread, rel_err(dpp * cgp, dpo * cgf))
        diff = rel_err(dpar, dpo * cgf)
        say("CG iterations device", int(cgres[0]), "oracle", ncg, "difference", diff, "oracle spread", spread)
        say("residuals device", cgres[1:int(cgres[0]) + 1], "oracle", cgro[1:ncg + 1])
        assert int(cgres[0]) == ncg
        assert diff < max(STOL, 30 * spread), (diff, spread)
        np.testing.assert_allclose(cgres[1:11], cgro[1:11], rtol=1e-8)


def test_headline_model_central_differences(crtmod):
    """The 240 x 80 model (19,521 nodes, 96 electrodes, 3,090 quadrupoles)
    with Dirichlet sides and bottom: 64 seeded cells, both directions."""
    import torch
    from titatest_amd import autograd as ad
    w = meshgen.large_grf(side_type=13)
    say("240x80 Dirichlet: adjoint gradient vs central differences on 64 cells")
    g = crtmod.CRT(w.mesh)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(1005)
    m0 = torch.tensor(np.log(w.sigma), dtype=torch.complex128, device=dev)
    N = w.nanz
    with torch.no_grad():
        d0 = ad.apparent_values(g, m0, w.strnr, w.vnr)
    dobs = d0 + torch.tensor(0.02 * rng.standard_normal(N) + 0.01j * rng.standard_normal(N),
                             dtype=torch.complex128, device=dev)
    wt = torch.tensor(1.0 / (0.02 + 1e-4 / np.abs(np.exp(-d0.cpu().numpy()))) ** 2, dtype=torch.float64,
                      device=dev)

    def misfit(m):
        r = dobs - ad.apparent_values(g, m, w.strnr, w.vnr)
        return torch.sum(wt * (r.real ** 2 + r.imag ** 2))

    m = m0.clone().requires_grad_(True)
    misfit(m).backward()
    grad = m.grad.detach().cpu().numpy()
    scale = np.max(np.abs(grad))
    h = 1e-6
    cells = rng.choice(m0.numel(), 64, replace=False)
    worst = 0.0
    for e in cells:
        for direction in (1.0, 1.0j):
            dm = torch.zeros_like(m0)
            dm[e] = h * direction
            with torch.no_grad():
                fd = (misfit(m0 + dm).item() - misfit(m0 - dm).item()) / (2 * h)
            an = grad[e].real if direction == 1.0 else grad[e].imag
            worst = max(worst, abs(fd - an) / scale)
            assert abs(fd - an) < 1e-6 * scale, (int(e), direction, fd, an)
    say("largest difference / largest gradient entry", worst)
