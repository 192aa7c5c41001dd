"""Uneven batches of independent models: one Gauss-Newton step each.

A step follows CRTomo's iteration for a fixed regularisation weight
(src/inv.f90 with src/update.f90:179-307): forward potentials and apparent
values d = -ln V (bvolti), the data-misfit gradient without forming J (the
adjoint kernels), the explicit Jacobian (bsensi), then the cgfac-preconditioned
Fletcher-Reeves CG of cjggra (src/cg_mod.f90:426-518) on J^H W J + lam R with
the rectangular smoothness operator (bsmatmreg) and right-hand side
J^H W (d_obs - d) - lam R m, m = ln sigma (update's bvec, src/update.f90:63-84
and :235-289).  The CG runs a fixed number of iterations (eps = 0).

Every model has its own handle, and so its own HIP stream; ``run_batch`` keeps
several of them in flight from a thread pool (the C-ABI calls release the
interpreter lock), so models of different mesh and electrode counts overlap on
the device.  ``use_graph`` runs each CG as one captured HIP graph
(crt_set_graph) instead of 4 kernel launches per iteration.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass

import numpy as np

from . import crt as _crt
from .meshgen import Workload, centroids

__all__ = ["GNModel", "prepare", "prepare_tri", "gn_step", "gn_iterate", "run_batch", "run_batch_persistent",
           "forward_many", "run_sequential", "rm_rect", "rm_tri"]


def rm_rect(smatm: np.ndarray, nx: int, m: np.ndarray) -> np.ndarray:
    """R m for the rectangular operator: diagonal smatm(:,1), right/left
    couplings smatm(:,2), down/up couplings smatm(:,3) (src/update.f90:63-84)."""
    M = len(m)
    out = smatm[:, 0] * m
    out[:-1] += smatm[:-1, 1] * m[1:]
    out[1:] += smatm[:-1, 1] * m[:-1]
    out[:-nx] += smatm[:-nx, 2] * m[nx:]
    out[nx:] += smatm[:-nx, 2] * m[:-nx]
    assert out.shape == (M,)
    return out


def rm_tri(smatm: np.ndarray, nachbar: np.ndarray, m: np.ndarray) -> np.ndarray:
    """R m for the triangular operator (ltri 1/2, src/update.f90:99-117): per
    element the neighbour terms smatm(i,k) m(nachbar(i,k)) in edge order, then
    the diagonal smatm(i, maxnodes+1) m(i).  nachbar (manz, maxnodes), 1-based,
    0 = no neighbour."""
    M, mx = nachbar.shape
    out = np.zeros(M, np.complex128)
    for k in range(mx):
        nb = nachbar[:, k]
        has = nb > 0
        out[has] = out[has] + smatm[has, k] * m[nb[has] - 1]
    return out + smatm[:, mx] * m


@dataclass
class GNModel:
    """One model of a batch, prepared once: handle, observed data, weights."""
    w: Workload
    g: "_crt.CRT"
    nx: int
    nz: int
    dobs: np.ndarray       # observed -ln V
    wgt: np.ndarray        # wmatd * wdfak
    smatm: np.ndarray      # bsmatmreg, manz x 3 (ltri 0), or bsmatmtri, manz x (maxnodes+1)
    rm: np.ndarray         # R ln(sigma)
    ltri: int = 0          # 0: rectangular stencil, 1: triangular operator with nachbar
    nachbar: np.ndarray = None  # (manz, maxnodes) edge neighbours for ltri 1

    def rm_of(self, m: np.ndarray) -> np.ndarray:
        return rm_rect(self.smatm, self.nx, m) if self.ltri == 0 else rm_tri(self.smatm, self.nachbar, m)

    def rdiag(self) -> np.ndarray:
        return self.smatm[:, 0] if self.ltri == 0 else self.smatm[:, -1]


def prepare(w: Workload, nx: int, nz: int, seed: int = 0, alfx: float = 1.0, alfz: float = 1.0,
            device_budget: int = 2 << 30, contrast: float = 1.2) -> GNModel:
    """Handle, wavenumbers, element matrices and the smoothness operator of a
    model on nx x nz quadrilaterals; synthetic observed data from a perturbed
    model (conductivity times ``contrast`` in a disc) with 2 % + 1e-4 Ohm
    Gaussian noise, and the error weights of CRTomo's error model
    (src/rdati.f90:254,315-320)."""
    g = _crt.CRT(w.mesh)
    g.set_device_budget(device_budget)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    smatm = g.smoothness(0, alfx, alfz, nx=nx, nz=nz)
    dobs, wgt = _observed(g, w, w.sigma, np.random.default_rng(seed), contrast)
    rm = rm_rect(smatm, nx, np.log(w.sigma))
    return GNModel(w, g, nx, nz, dobs, wgt, smatm, rm)


def _observed(g, w, sigma, rng, contrast):
    """Synthetic observed -ln V from a model perturbed by ``contrast`` in a
    disc, with 2 % + 1e-4 Ohm Gaussian noise, and CRTomo's error-model weights
    (src/rdati.f90:254,315-320)."""
    cx, cz = centroids(w.mesh)
    x0 = rng.uniform(cx.min(), cx.max())
    z0 = rng.uniform(cz.min() * 0.5, 0.0)
    blob = (cx - x0) ** 2 + (cz - z0) ** 2 < (0.2 * (cx.max() - cx.min())) ** 2
    g.forward(sigma * np.where(blob, contrast, 1.0))
    v, _ = g.bvolti(w.strnr, w.vnr)
    r = np.abs(v) * (1.0 + 0.02 * rng.standard_normal(len(v))) + 1e-4 * rng.standard_normal(len(v))
    dobs = -np.log(np.abs(r)) - 1j * np.angle(v)
    wgt = 1.0 / (0.02 + 1e-4 / np.abs(v)) ** 2
    return dobs, wgt


def prepare_tri(w: Workload, seed: int = 0, alfx: float = 1.0, alfz: float = 1.0, lsytop: bool = True,
                contrast: float = 1.2) -> GNModel:
    """As ``prepare`` for a triangular (or any unstructured) mesh: the
    triangular smoothness operator of bsmatmtri with bnachbar's neighbours
    (ltri = 1)."""
    g = _crt.CRT(w.mesh, lsytop=lsytop)
    g.rwaven(w.strnr, w.vnr)
    g.precal()
    smatm, nb = g.smoothness(1, alfx, alfz)
    mx = w.mesh.maxnodes
    nachbar = np.ascontiguousarray(nb[:, :mx])
    dobs, wgt = _observed(g, w, w.sigma, np.random.default_rng(seed), contrast)
    rm = rm_tri(smatm, nachbar, np.log(w.sigma))
    return GNModel(w, g, 0, 0, dobs, wgt, smatm, rm, ltri=1, nachbar=nachbar)


def gn_step(mod: GNModel, lam: float = 10.0, ncg: int = 20, use_graph: bool = False):
    """One Gauss-Newton step on the model's current conductivities: returns
    (dpar, grad, cgres), dpar = the scaled model update, grad = the adjoint
    gradient J^H W (d_obs - d)."""
    g, w = mod.g, mod.w
    g.set_graph(use_graph)
    g.forward(w.sigma)
    _, sa = g.bvolti(w.strnr, w.vnr)
    resid = mod.dobs - sa
    grad = g.adjoint_grad(mod.wgt, resid, mode=1)
    g.sensitivities(1, to_host=False)
    dpar, _, cgres = g.cg_solve(mod.wgt, mod.dobs, sa, lam, 0, mod.smatm, mod.smatm[:, 0], ncg, eps=0.0,
                                nx=mod.nx, rm=mod.rm)
    return dpar, grad, cgres


def gn_iterate(mod: GNModel, iters: int = 3, lam: float = 10.0, ncg: int = 20, step: float = 1.0,
               eps: float = 0.0, lphi0: bool = False, use_graph: bool = False, sigma0=None):
    """Fixed-lambda Gauss-Newton iterations in the order of src/inv.f90: the
    forward and -ln V of the current model, the data misfit (dmisft,
    :524/:677), the explicit Jacobian, the CG update of update (right-hand
    side J^H W (d_obs - d) - lam R m with m = ln sigma), then bpar, the model
    update par + step dpar and bsigma (:924-933).  Starts from ``sigma0``
    (default: the model's conductivities, left unchanged); there is no line
    search (the reference halves a step that raises the misfit and fits a
    parabola to the step length, :872-912),
    so lam must keep the steps in the linear range.  Returns (misfits, sigma) with
    misfits[i] = (nrmsd, betrms, pharms) of the model before iteration i and
    the final entry that of the last model."""
    g, w = mod.g, mod.w
    g.set_graph(use_graph)
    g.set_model_map()
    sigma = np.array(w.sigma if sigma0 is None else sigma0, np.complex128)
    hist = []
    for it in range(iters + 1):
        g.forward(sigma)
        _, sa = g.bvolti(w.strnr, w.vnr)
        st, _, _ = g.dmisft(mod.dobs, sa, mod.wgt, mod.wgt, mod.wgt)
        hist.append(tuple(st[:3]))
        if it == iters:
            break
        g.sensitivities(1, to_host=False)
        rm = mod.rm_of(np.log(sigma))
        dpar, _, _ = g.cg_solve(mod.wgt, mod.dobs, sa, lam, mod.ltri, mod.smatm, mod.rdiag(), ncg, eps=eps,
                                nx=mod.nx, nachbar=mod.nachbar, rm=rm)
        par = g.bpar(sigma)
        par, _, _ = g.update_step(par, dpar, step, lphi0)
        sigma = g.bsigma(par)
    return hist, sigma


def run_batch(models, threads: int = 8, **kw):
    """Steps of all models, ``threads`` of them in flight; returns (results,
    wall seconds)."""
    t = time.perf_counter()
    with ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(lambda m: gn_step(m, **kw), models))
    return res, time.perf_counter() - t


def forward_many(handles, sigmas):
    """One crt_forward_many over several models: ``handles`` CRT objects,
    ``sigmas`` their complex conductivities (host arrays); the conductivities
    are staged in device memory for the call."""
    import torch
    dev = torch.device("cuda", torch.cuda.current_device())
    ts = [torch.as_tensor(np.ascontiguousarray(s, np.complex128), device=dev) for s in sigmas]
    torch.cuda.synchronize()
    _crt.forward_many(handles, [t.data_ptr() for t in ts])
    return ts


def run_batch_persistent(models, threads: int = 8, lam: float = 10.0, ncg: int = 20, use_graph: bool = True):
    """Gauss-Newton steps of all models with the forwards batched: one
    crt_forward_many over every model (persistent factor launches fed from a
    work queue across models), then per model -ln V, the adjoint gradient,
    the explicit Jacobian and the CG, ``threads`` models in flight.  Returns
    (results, wall seconds); the results equal gn_step's bit for bit."""
    t = time.perf_counter()
    forward_many([m.g for m in models], [m.w.sigma for m in models])

    def rest(mod):
        g, w = mod.g, mod.w
        g.set_graph(use_graph)
        _, sa = g.bvolti(w.strnr, w.vnr)
        grad = g.adjoint_grad(mod.wgt, mod.dobs - sa, mode=1)
        g.sensitivities(1, to_host=False)
        dpar, _, cgres = g.cg_solve(mod.wgt, mod.dobs, sa, lam, mod.ltri, mod.smatm, mod.rdiag(), ncg, eps=0.0,
                                    nx=mod.nx, nachbar=mod.nachbar, rm=mod.rm)
        return dpar, grad, cgres

    with ThreadPoolExecutor(max_workers=threads) as ex:
        res = list(ex.map(rest, models))
    return res, time.perf_counter() - t


def run_sequential(models, lam: float = 10.0, ncg: int = 20):
    """The plain per-model path the batched one is measured against: one model
    at a time, forward, -ln V, the explicit Jacobian and J^H W r (the data
    term from J, as CRTomo forms it), then the CG with launch-by-launch
    iterations.  Returns (results, wall seconds)."""
    t = time.perf_counter()
    res = []
    for mod in models:
        g, w = mod.g, mod.w
        g.set_graph(False)
        g.forward(w.sigma)
        _, sa = g.bvolti(w.strnr, w.vnr)
        g.sensitivities(1, to_host=False)
        grad = g.jhw_r(mod.wgt, mod.dobs - sa)
        dpar, _, cgres = g.cg_solve(mod.wgt, mod.dobs, sa, lam, mod.ltri, mod.smatm, mod.rdiag(), ncg, eps=0.0,
                                    nx=mod.nx, nachbar=mod.nachbar, rm=mod.rm)
        res.append((dpar, grad, cgres))
    return res, time.perf_counter() - t
