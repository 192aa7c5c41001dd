"""torch.autograd binding of the forward path and its adjoint gradient.

``apparent_values(crt, m, strnr, vnr)`` maps the log complex conductivities
m = ln(sigma) (complex128 CUDA tensor, one per model cell) to CRTomo's data,
d = -ln V (src/bvolti.f90:88, complex128 CUDA tensor, one per quadrupole), with
the whole forward (assembly, band factorisation, substitutions, back-transform,
voltage pick-off) running in the HIP kernels behind the C-ABI.

Its backward is the adjoint gradient: for an incoming gradient g it returns
J^H g with J = d(-ln V)/d(ln sigma) as bsensi defines it
(src/bsensi.f90:23-194), computed from the stored potentials without forming
J (crt_adjoint_gradient_dev).  PyTorch's convention for complex inputs (the
gradient of a real loss is dL/d(Re m) + i dL/d(Im m), and for a holomorphic map
the vector-Jacobian product is J^H g) makes that exactly the backward of the
holomorphic map m -> d.  As in the reference, J leaves out the conductivity
dependence of the mixed-boundary terms (src/bsens.f90:11); on meshes without
mixed boundaries it is the exact derivative.

The Function fails loudly without the HIP library or a device: it has no CPU
path.
"""
from __future__ import annotations

import numpy as np
import torch

from . import crt as _crt

__all__ = ["ApparentValues", "apparent_values"]


def _check(where, rc):
    if rc:
        raise _crt.CRTError(where, rc)


class ApparentValues(torch.autograd.Function):
    """m = ln(sigma) (cells) -> d = -ln V (quadrupoles), both complex128 on the
    device the handle runs on."""

    @staticmethod
    def forward(ctx, m, g: "_crt.CRT", strnr, vnr):
        if not m.is_cuda or m.dtype != torch.complex128:
            raise TypeError("m must be a complex128 CUDA tensor")
        st = np.ascontiguousarray(strnr, np.int32)
        vn = np.ascontiguousarray(vnr, np.int32)
        n_q = len(st)
        sigma = torch.exp(m.detach()).contiguous()
        volt = torch.empty(n_q, dtype=torch.complex128, device=m.device)
        d = torch.empty(n_q, dtype=torch.complex128, device=m.device)
        # the C-ABI runs on its own stream and returns synchronised; the
        # inputs come from torch's stream
        torch.cuda.current_stream(m.device).synchronize()
        _check("forward_dev", g.lib.crt_forward_dev(g.h, sigma.data_ptr()))
        _check("voltages_dev", g.lib.crt_voltages_dev(g.h, n_q, st.ctypes.data, vn.ctypes.data, 1,
                                                      volt.data_ptr(), d.data_ptr()))
        g.nanz = n_q
        # the handle keeps the potentials and voltages of its last forward;
        # a later forward on the same handle makes backward recompute them
        g._ad_generation = getattr(g, "_ad_generation", 0) + 1
        ctx.generation = g._ad_generation
        ctx.g = g
        ctx.n_cells = m.numel()
        ctx.save_for_backward(sigma)
        ctx.st, ctx.vn = st, vn
        return d

    @staticmethod
    def backward(ctx, grad_d):
        g = ctx.g
        if getattr(g, "_ad_generation", 0) != ctx.generation:
            (sigma,) = ctx.saved_tensors
            n_q = len(ctx.st)
            scratch = torch.empty((2, n_q), dtype=torch.complex128, device=sigma.device)
            torch.cuda.current_stream(sigma.device).synchronize()
            _check("forward_dev", g.lib.crt_forward_dev(g.h, sigma.data_ptr()))
            _check("voltages_dev", g.lib.crt_voltages_dev(g.h, n_q, ctx.st.ctypes.data, ctx.vn.ctypes.data,
                                                          1, scratch[0].data_ptr(), scratch[1].data_ptr()))
            g._ad_generation = ctx.generation
        grad_d = grad_d.to(torch.complex128).contiguous()
        ones = torch.ones(grad_d.numel(), dtype=torch.float64, device=grad_d.device)
        grad_m = torch.empty(ctx.n_cells, dtype=torch.complex128, device=grad_d.device)
        torch.cuda.current_stream(grad_d.device).synchronize()
        _check("adjoint_gradient_dev", g.lib.crt_adjoint_gradient_dev(
            g.h, 1, ones.data_ptr(), grad_d.data_ptr(), grad_m.data_ptr()))
        return grad_m, None, None, None


def apparent_values(g: "_crt.CRT", m: torch.Tensor, strnr, vnr) -> torch.Tensor:
    """-ln V of the quadrupoles (strnr, vnr) for log conductivities m; see
    ApparentValues.  The handle must have its wavenumbers and precal set."""
    return ApparentValues.apply(m, g, strnr, vnr)
