"""torch.autograd.Function wrappers of the hand-written adjoint.

The backward pass of the force evaluation is the HIP adjoint kernel itself:
forces = -dE/dx come out of the forward call (dE/dB -> u -> per-pair
recomputation of the radial basis and Y_lm), so backward only scales them by
the incoming gradient.  dE/dbeta = sum_i d_i d_i^T per species (d = B/|B|).
"""

import torch

from . import _capi


class MappedEnergy(torch.autograd.Function):
    """E, virial = MappedEnergy.apply(positions, beta, pair, cell, species, form)

    positions (n, 3) float64 on the GPU, beta the unpacked coefficients
    (S, F, F) float64, pair a :class:`PairFLARE`, species (n,) int32 on the
    GPU.  Gradients flow to positions and beta; the virial (xx yy zz xy xz
    yz) is returned as a non-differentiable output.  Only the symmetric part
    of beta enters E = d^T beta d, and the force kernel assumes a symmetric
    beta, so (beta + beta^T) / 2 is evaluated; dE/dbeta = d d^T holds for
    any beta."""

    @staticmethod
    def forward(ctx, positions, beta, pair, cell, species, form):
        pos = positions.detach().contiguous()
        b = beta.detach()
        if b.dim() == 3:
            b = 0.5 * (b + b.transpose(1, 2))
        b = b.contiguous()
        n = pos.shape[0]
        dev = pos.device
        pair.build_neighbors(cell, pos, species)
        e = torch.empty(1, dtype=torch.float64, device=dev)
        f = torch.empty((max(n, 1), 3), dtype=torch.float64, device=dev)
        v = torch.empty(6, dtype=torch.float64, device=dev)
        pair.compute_device(species, form=form, d_beta=b, energy=e, forces=f,
                            virial=v)
        gb = torch.empty_like(b)
        pair.beta_grad_device(gb)
        ctx.save_for_backward(f[:n], gb)
        ctx.mark_non_differentiable(v)
        return e[0], v

    @staticmethod
    def backward(ctx, g_e, g_v):
        f, gb = ctx.saved_tensors
        return -f * g_e, gb * g_e, None, None, None, None


def mapped_energy(positions, beta, pair, cell, species, form="adjoint_mixed"):
    return MappedEnergy.apply(positions, beta, pair, cell, species, form)


class GPMeanEnergy(torch.autograd.Function):
    """E, stress = GPMeanEnergy.apply(positions, gp, structure, form)

    Energy of SparseGP.predict_mean with the forces of the adjoint kernel as
    -dE/dx; `structure` supplies cell, species and the B2 calculator and has
    its positions replaced by `positions` before the call."""

    @staticmethod
    def forward(ctx, positions, gp, structure, form):
        pos = positions.detach().contiguous()
        structure.d_positions = pos
        structure.positions = pos.cpu().numpy()
        structure.compute_neighbors()
        efs = torch.as_tensor(gp.predict_mean(structure, form=form),
                              device=pos.device)
        n = pos.shape[0]
        forces = efs[1:1 + 3 * n].reshape(n, 3)
        stress = efs[1 + 3 * n:]
        ctx.save_for_backward(forces)
        ctx.mark_non_differentiable(stress)
        return efs[0], stress

    @staticmethod
    def backward(ctx, g_e, g_s):
        (f,) = ctx.saved_tensors
        return -f * g_e, None, None, None


def gp_mean_energy(positions, gp, structure, form="adjoint_fp64"):
    return GPMeanEnergy.apply(positions, gp, structure, form)


__all__ = ["MappedEnergy", "mapped_energy", "GPMeanEnergy", "gp_mean_energy",
           "_capi"]
