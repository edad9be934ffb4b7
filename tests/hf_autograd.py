"""torch.autograd wrapper of the Hellmann-Feynman gradient, for the tests only.

``free_energy(theta, spec)`` returns f = -ln(lambda0 / normfactor) / L for
theta = (betaJ, kappa, kappa2, h) as a float64 CPU tensor; its backward pass
calls ``tm_grad_hf`` (dual-number operator kernels on the GPU) with the
solver's right and left eigenvectors.  torch is only the autograd machinery
here: the eigensolve and the derivative run through the C ABI.
"""
import numpy as np
import torch


class LeadingLogEigenvalue(torch.autograd.Function):
    """ln(lambda0 / normfactor) of the transfer matrix at parameters theta."""

    @staticmethod
    def forward(ctx, theta, spec):
        import latttmat_amd as lt
        bJ, k, k2, h = (float(v) for v in theta.detach().cpu().numpy())
        op = lt.TransferMatrix(spec["model"], spec.get("dim", 2), spec["axialdir"], spec["size"],
                               J=bJ, kappa=k, kappa2=k2, h=h, neigs=1)
        sym = op.is_symmetric()
        r = lt.solve(op, nev=1, tol=1e-12, want_vectors=True, want_left=not sym)
        lam = float(np.real(r.evals[0]))
        right = np.real(r.vecs[:, 0])
        left = right if sym else r.vecl
        dlam, dlnnf = op.grad_hf(left, right)
        ctx.save_for_backward(torch.as_tensor(dlam / lam - dlnnf, dtype=torch.float64))
        return torch.tensor(np.log(lam / op.getnormfactor()), dtype=torch.float64)

    @staticmethod
    def backward(ctx, g):
        (dlog,) = ctx.saved_tensors
        return g * dlog, None


def free_energy(theta, spec):
    return -LeadingLogEigenvalue.apply(theta, spec) / spec["layer"]
