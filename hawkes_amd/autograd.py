"""torch.autograd wrapper of the engine's log-likelihood and exact gradient.

``loglik(params, engine)`` returns a 0-d tensor whose backward pass applies the gradient the
engine computed in the same pair pass (no finite differences). ``loglik_batch(params, engine)``
does the same for a K x 6 tensor of parameter vectors in one batched launch. torch is used only
for the tensor plumbing; the computation is the HIP engine's.
"""
import numpy as np
import torch


class HawkesLogLikelihood(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, engine):
        p = params.detach().to("cpu", torch.float64).numpy()
        if p.shape != (6,):
            raise ValueError("params must have shape (6,)")
        engine.set_parameters(p)
        ll, g = engine.log_likelihood_and_gradient()
        ctx.save_for_backward(torch.as_tensor(g, dtype=params.dtype, device=params.device))
        return torch.tensor(ll, dtype=params.dtype, device=params.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out * g, None


class HawkesLogLikelihoodBatch(torch.autograd.Function):
    @staticmethod
    def forward(ctx, params, engine):
        p = params.detach().to("cpu", torch.float64).numpy()
        if p.ndim != 2 or p.shape[1] != 6:
            raise ValueError("params must have shape (K, 6)")
        engine.set_parameters_batch(p)
        ll, g = engine.log_likelihood_batch(gradient=True)
        ctx.save_for_backward(torch.as_tensor(g, dtype=params.dtype, device=params.device))
        return torch.as_tensor(ll, dtype=params.dtype, device=params.device)

    @staticmethod
    def backward(ctx, grad_out):
        (g,) = ctx.saved_tensors
        return grad_out[:, None] * g, None


def loglik(params, engine):
    """Log-likelihood at params = (sigmaXprec, tauXprec, tauTprec, omega, theta, mu0)."""
    return HawkesLogLikelihood.apply(params, engine)


def loglik_batch(params, engine):
    """Log-likelihoods of K parameter vectors (K x 6) from one batched launch."""
    return HawkesLogLikelihoodBatch.apply(params, engine)


def as_numpy(x):
    return np.asarray(x.detach().cpu().numpy() if isinstance(x, torch.Tensor) else x, dtype=np.float64)
