"""PyTorch's vector-Jacobian product for complex holomorphic maps is J^H g.

titatest_amd/autograd.py returns J^H g (the adjoint gradient) as the backward
of m -> -ln V; this pins the convention it relies on with a holomorphic map
whose derivative is known in closed form.  CPU only.
"""
import numpy as np
import torch


def test_complex_backward_is_conjugate_jacobian_product():
    rng = np.random.default_rng(3)
    z = torch.tensor(rng.standard_normal(5) + 1j * rng.standard_normal(5), dtype=torch.complex128,
                     requires_grad=True)
    c = torch.tensor(0.3 - 1.7j, dtype=torch.complex128)
    f = c * z + z * z            # f'(z) = c + 2 z, elementwise
    g = torch.tensor(rng.standard_normal(5) + 1j * rng.standard_normal(5), dtype=torch.complex128)
    f.backward(g)
    deriv = (c + 2 * z).detach()
    np.testing.assert_allclose(z.grad.numpy(), (deriv.conj() * g).numpy(), rtol=1e-14)


def test_real_loss_gradient_is_real_plus_i_imag_partials():
    z0 = np.array([0.4 + 0.9j, -1.2 + 0.1j])
    z = torch.tensor(z0, dtype=torch.complex128, requires_grad=True)
    loss = torch.sum((z * z).real ** 2 + (3 * z).imag ** 2)
    loss.backward()
    h = 1e-6
    for e in range(2):
        for d in (1.0, 1.0j):
            zp, zm = z0.copy(), z0.copy()
            zp[e] += h * d
            zm[e] -= h * d
            L = lambda v: np.sum(((v * v).real) ** 2 + ((3 * v).imag) ** 2)
            fd = (L(zp) - L(zm)) / (2 * h)
            an = z.grad[e].real.item() if d == 1.0 else z.grad[e].imag.item()
            assert abs(fd - an) < 1e-6
