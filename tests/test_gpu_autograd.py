"""torch.autograd wrappers of the adjoint kernel: gradients against central
finite differences (h = 1e-4 A) and dE/dbeta."""
import numpy as np
import pytest
import torch

from flare_pp_amd import PairFLARE, pack_beta, synth
from flare_pp_amd.autograd import mapped_energy

pytestmark = pytest.mark.gpu


def test_mapped_energy_gradients():
    S, N, L = 2, 4, 3
    cell, species, pos = synth.fcc(2, 4.05, 0.08, [16, 16], seed=4)
    F = (S * N) * (S * N + 1) // 2 * (L + 1)
    beta = synth.symmetric_beta(S, F, 1e-2, seed=5)
    pair = PairFLARE.from_coefficients(2, S, N, L, np.full((S, S), 5.0),
                                       np.stack([pack_beta(b) for b in beta]))
    dev = "cuda:0"
    x = torch.tensor(pos, device=dev, requires_grad=True)
    b = torch.tensor(beta, device=dev, requires_grad=True)
    spc = torch.tensor(species, device=dev, dtype=torch.int32)
    for form in ("adjoint_fp64", "plain"):
        E, vir = mapped_energy(x, b, pair, cell, spc, form)
        gx, gb = torch.autograd.grad(E, (x, b))
        h = 1e-4
        rng = np.random.default_rng(0)
        for a in rng.choice(len(pos), 6, replace=False):
            for c in range(3):
                xp = x.detach().clone()
                xm = x.detach().clone()
                xp[a, c] += h
                xm[a, c] -= h
                fd = (mapped_energy(xp, b, pair, cell, spc, form)[0]
                      - mapped_energy(xm, b, pair, cell, spc, form)[0]) / (2 * h)
                assert abs(fd.item() - gx[a, c].item()) < 1e-6 * max(1, gx.abs().max().item())
        # dE/dbeta = sum_i d_i d_i^T: E is linear in beta
        E0 = E.item()
        d = torch.zeros_like(b)
        d[1, 3, 7] = d[1, 7, 3] = 0.5
        E1 = mapped_energy(x, b + d, pair, cell, spc, form)[0].item()
        assert abs((E1 - E0) - (gb * d).sum().item()) < 1e-10 * max(1, abs(E0))
        assert torch.allclose(gb, gb.transpose(1, 2))
