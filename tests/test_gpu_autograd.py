"""torch.autograd binding (titatest_amd/autograd.py): the adjoint gradient as
the backward of m = ln(sigma) -> d = -ln V.

* central finite differences of a weighted misfit on a mesh without mixed
  boundaries (the reference's sensitivity is the exact derivative there),
  step 1e-6 in the real and the imaginary direction of ln(sigma), agreement
  to 1e-6 relative (the differences' own truncation and round-off);
* J^H g from the explicit Jacobian (crt_sensitivities, bsensi semantics) on a
  mesh with mixed boundaries, to the sensitivity tolerance 1e-8;
* a backward after another forward on the same handle recomputes its own
  potentials.
"""
import numpy as np
import pytest
import torch

from conftest import rel_err
from titatest_amd import meshgen

pytestmark = pytest.mark.gpu

STOL = 1e-8


def say(*a):
    print("[gpu-case]", *a, flush=True)


@pytest.fixture(scope="module")
def ad(hip_lib):
    from titatest_amd import autograd, crt
    return crt, autograd


def _setup(crt, w):
    g = crt.CRT(w.mesh)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    return g


def _misfit(ad_mod, g, w, m, dobs, wt):
    d = ad_mod.apparent_values(g, m, w.strnr, w.vnr)
    r = dobs - d
    return torch.sum(wt * (r.real ** 2 + r.imag ** 2))


def test_autograd_matches_central_differences(ad):
    crt, ad_mod = ad
    w = meshgen.small_dirichlet()
    say("small_dirichlet autograd vs central differences")
    g = _setup(crt, w)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(2024)
    m0 = torch.tensor(np.log(w.sigma), dtype=torch.complex128, device=dev)
    with torch.no_grad():
        d0 = ad_mod.apparent_values(g, m0, w.strnr, w.vnr)
    dobs = d0 + torch.tensor(0.02 * rng.standard_normal(len(w.strnr)), dtype=torch.complex128, device=dev)
    wt = torch.tensor(rng.uniform(0.5, 2.0, len(w.strnr)), dtype=torch.float64, device=dev)
    m = m0.clone().requires_grad_(True)
    loss = _misfit(ad_mod, g, w, m, dobs, wt)
    loss.backward()
    grad = m.grad.detach().cpu().numpy()
    h = 1e-6
    cells = rng.choice(m0.numel(), 8, replace=False)
    for e in cells:
        for direction in (1.0, 1.0j):
            dm = torch.zeros_like(m0)
            dm[e] = h * direction
            with torch.no_grad():
                lp = _misfit(ad_mod, g, w, m0 + dm, dobs, wt).item()
                lm = _misfit(ad_mod, g, w, m0 - dm, dobs, wt).item()
            fd = (lp - lm) / (2 * h)
            an = grad[e].real if direction == 1.0 else grad[e].imag
            scale = np.max(np.abs(grad))
            assert abs(fd - an) < 1e-6 * scale, (e, direction, fd, an)


def test_autograd_equals_explicit_jacobian(ad):
    crt, ad_mod = ad
    w = meshgen.small_block()
    say("small_block autograd vs J^H g")
    g = _setup(crt, w)
    dev = torch.device("cuda:0")
    rng = np.random.default_rng(7)
    m = torch.tensor(np.log(w.sigma), dtype=torch.complex128, device=dev, requires_grad=True)
    d = ad_mod.apparent_values(g, m, w.strnr, w.vnr)
    gout = torch.tensor(rng.standard_normal(len(w.strnr)) + 1j * rng.standard_normal(len(w.strnr)),
                        dtype=torch.complex128, device=dev)
    d.backward(gout)
    grad = m.grad.detach().cpu().numpy()
    J = g.bsensi()  # N x M, d(-ln V)/d(ln sigma)
    ref = J.conj().T @ gout.cpu().numpy()
    assert rel_err(grad, ref) < STOL


def test_autograd_backward_after_other_forward(ad):
    crt, ad_mod = ad
    w = meshgen.small_block()
    say("small_block backward after a second forward")
    g = _setup(crt, w)
    dev = torch.device("cuda:0")
    m1 = torch.tensor(np.log(w.sigma), dtype=torch.complex128, device=dev, requires_grad=True)
    d1 = ad_mod.apparent_values(g, m1, w.strnr, w.vnr)
    gout = torch.ones_like(d1)
    d1.backward(gout, retain_graph=True)
    ref = m1.grad.detach().clone()
    m1.grad = None
    m2 = (m1.detach() + 0.3).requires_grad_(True)
    ad_mod.apparent_values(g, m2, w.strnr, w.vnr)   # the handle now holds m2's fields
    d1.backward(gout)
    assert rel_err(m1.grad.detach().cpu().numpy(), ref.cpu().numpy()) < 1e-14
