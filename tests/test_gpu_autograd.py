"""torch.autograd wrappers of the adjoint kernel: gradients against central
finite differences (h = 1e-4 A) and dE/dbeta."""
import numpy as np
import pytest
import torch

from flare_pp_amd import PairFLARE, _capi, pack_beta, synth
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


def bench_settings_pair(seed):
    S, N, L = 2, 8, 4
    F = (S * N) * (S * N + 1) // 2 * (L + 1)
    beta = synth.symmetric_beta(S, F, 1e-3, seed=seed)
    return PairFLARE.from_coefficients(2, S, N, L, np.full((S, S), 5.0),
                                       np.stack([pack_beta(b) for b in beta])), F


def beta_grad(pair, F):
    g = torch.empty((pair.n_species, F, F), dtype=torch.float64, device="cuda:0")
    pair.beta_grad_device(g)
    return g.cpu().numpy()


def test_beta_grad_forms_and_ownership():
    """dE/dbeta = sum_i d_i d_i^T after an FP32 or mixed evaluation (from the
    FP32 rows B/|B| and from the BF16 hi + lo rows) against the FP64 one; a
    later evaluation of another pair on the same context invalidates the
    first pair's gradient rather than handing back the other pair's rows."""
    cell, species, pos = synth.fcc(3, 4.05, 0.08, [54, 54], seed=31)
    a, F = bench_settings_pair(32)
    b, _ = bench_settings_pair(33)
    a.compute(cell, species, pos, form="adjoint_fp64")
    g64 = beta_grad(a, F)
    scale = np.abs(g64).max()
    # exact: the same rows give the same gradient from the plain form
    a.compute(cell, species, pos, form="plain")
    assert np.array_equal(beta_grad(a, F), g64)
    # FP32 rows: ~1e-7 relative; BF16 hi + lo rows: 16 significand bits each
    for form, tol in (("adjoint_fp32", 1e-6), ("adjoint_mixed", 1e-4)):
        a.compute(cell, species, pos, form=form)
        err = np.abs(beta_grad(a, F) - g64).max() / scale
        print(f"beta_grad {form}: max rel err {err:.2e}", flush=True)
        assert err < tol, form
    # pair b (other species layout of rows, same sizes) overwrites the rows
    cell2, species2, pos2 = synth.fcc(2, 4.05, 0.08, [16, 16], seed=34)
    a.compute(cell, species, pos, form="adjoint_fp64")
    b.compute(cell2, species2, pos2, form="adjoint_fp64")
    ctx = _capi.context(0)
    g = torch.empty((2, F, F), dtype=torch.float64, device="cuda:0")
    rc = _capi.lib().flare_pair_beta_grad(ctx.handle, a._h, _capi.dptr(g))
    assert rc == 1 and b"overwritten" in _capi.lib().flare_last_error()
    gb = beta_grad(b, F)  # b's own rows are still current
    a.compute(cell, species, pos, form="adjoint_fp64")
    assert np.array_equal(beta_grad(a, F), g64)
    assert not np.allclose(gb, g64)


def test_mapped_energy_beta_direction_low_precision():
    """E is linear in beta: E(beta + D) - E(beta) = <dE/dbeta, D> in the
    FP32 and mixed forms (D symmetric), to the forms' energy precision."""
    cell, species, pos = synth.fcc(3, 4.05, 0.08, [54, 54], seed=35)
    pair, F = bench_settings_pair(36)
    dev = "cuda:0"
    x = torch.tensor(pos, device=dev)
    b = torch.tensor(pair.beta, device=dev, requires_grad=True)
    spc = torch.tensor(species, device=dev, dtype=torch.int32)
    rng = np.random.default_rng(37)
    D = rng.normal(0, 1e-3, (2, F, F))
    D = torch.tensor(0.5 * (D + D.transpose(0, 2, 1)), device=dev)
    for form, tol in (("adjoint_fp64", 1e-10), ("adjoint_fp32", 1e-5),
                      ("adjoint_mixed", 1e-4)):
        E0, _ = mapped_energy(x, b, pair, cell, spc, form)
        (gb,) = torch.autograd.grad(E0, (b,))
        E1 = mapped_energy(x, b + D, pair, cell, spc, form)[0].item()
        lin = (gb * D).sum().item()
        err = abs((E1 - E0.item()) - lin) / abs(lin)
        print(f"dE/dbeta direction {form}: rel err {err:.2e}", flush=True)
        assert err < tol, form


def test_mapped_energy_symmetrizes_beta():
    """Only the symmetric part of beta enters E = d^T beta d: an asymmetric
    beta gives the energy and forces of (beta + beta^T) / 2, and dE/dbeta =
    d d^T still holds for it."""
    cell, species, pos = synth.fcc(2, 4.05, 0.08, [16, 16], seed=38)
    pair, F = bench_settings_pair(39)
    dev = "cuda:0"
    x = torch.tensor(pos, device=dev, requires_grad=True)
    spc = torch.tensor(species, device=dev, dtype=torch.int32)
    bs = torch.tensor(pair.beta, device=dev)
    A = torch.tensor(np.random.default_rng(40).normal(0, 1e-3, (2, F, F)), device=dev)
    asym = (bs + A - A.transpose(1, 2)).requires_grad_(True)
    for form in ("adjoint_fp64", "adjoint_fp32"):
        E, v = mapped_energy(x, asym, pair, cell, spc, form)
        gx, gb = torch.autograd.grad(E, (x, asym))
        Es, vs = mapped_energy(x, bs, pair, cell, spc, form)
        (gxs,) = torch.autograd.grad(Es, (x,))
        assert torch.allclose(gx, gxs, rtol=0, atol=1e-10 * gxs.abs().max().item()), form
        assert abs(E.item() - Es.item()) < 1e-10 * abs(Es.item()), form
        assert torch.allclose(gb, gb.transpose(1, 2))
