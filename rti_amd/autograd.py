"""torch.autograd binding of the flux trace (thin test binding).

``FluxFunction.apply(s, n, scene, params)`` returns the per-element
deposited flux (the reference's ``value``: sum of deposits, in input order)
as a float64 tensor; its backward uses d(value)/ds and d(value)/dn from the
adjoint accumulators of the same GPU run (``rti_get_adjoint``):
grad_s = sum_i grad_i * d value_i / ds, likewise for n.
"""
import numpy as np
import torch

from ._lib import Context


class FluxFunction(torch.autograd.Function):
    @staticmethod
    def forward(ctx, s, n, scene, params):
        kw = dict(params)
        kw.update(sticking=float(s.item()), cos_exponent=float(n.item()), adjoint=1)
        with Context() as c:
            c.set_scene(scene)
            c.set_params(**kw)
            c.run()
            raw = c.flux()["raw"]
            ds, dn = c.adjoint()
        ctx.save_for_backward(torch.from_numpy(ds), torch.from_numpy(dn))
        return torch.from_numpy(np.asarray(raw, dtype=np.float64))

    @staticmethod
    def backward(ctx, grad):
        ds, dn = ctx.saved_tensors
        g = grad.to(torch.float64)
        return (g * ds).sum().reshape(1), (g * dn).sum().reshape(1), None, None


def flux(s, n, scene, **params):
    """Deposited flux per element as a differentiable function of s and n."""
    return FluxFunction.apply(s, n, scene, params)
