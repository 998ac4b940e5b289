"""flare_pair_compute_noalloc (workspace in the device's static scratch)
against flare_pair_compute on the same neighbor list: identical results in
every form, on two calls in a row with the second input smaller, dE/dbeta
from the scratch rows, and a clear refusal of an input too large for the
scratch."""
import ctypes

import numpy as np
import pytest
import torch

from flare_pp_amd import PairFLARE, _capi, pack_beta, synth

pytestmark = pytest.mark.gpu

FORMS = ("plain", "adjoint_fp64", "adjoint_fp32", "adjoint_mixed")
DEV = "cuda:0"


def bench_pair():
    S, N, L = 2, 8, 4
    F = (S * N) * (S * N + 1) // 2 * (L + 1)
    beta = synth.symmetric_beta(S, F, 1e-3, seed=11)
    packed = np.stack([pack_beta(b) for b in beta])
    return PairFLARE.from_coefficients(2, S, N, L, np.full((S, S), 5.0), packed)


def run(pair, spc, form, noalloc):
    n = spc.numel()
    e = torch.full((1,), np.nan, dtype=torch.float64, device=DEV)
    f = torch.full((n, 3), np.nan, dtype=torch.float64, device=DEV)
    v = torch.full((6,), np.nan, dtype=torch.float64, device=DEV)
    ea = torch.full((n,), np.nan, dtype=torch.float64, device=DEV)
    pair.compute_device(spc, form=form, energy=e, forces=f, virial=v, eatom=ea,
                        noalloc=noalloc)
    torch.cuda.synchronize()
    return [t.cpu().numpy() for t in (e, f, v, ea)]


def structure(n_cells, seed):
    cell, species, pos = synth.fcc(n_cells, 4.05, 0.08,
                                   [2 * n_cells ** 3, 2 * n_cells ** 3], seed)
    return (cell, torch.tensor(pos, device=DEV),
            torch.tensor(species, dtype=torch.int32, device=DEV))


def test_noalloc_matches_compute_two_calls():
    pair = bench_pair()
    F = pair.n_descriptors
    # 4 x 4 x 4 cells (256 atoms), then 3 x 3 x 3 (108 atoms)
    for n_cells, seed in ((4, 21), (3, 22)):
        cell, pos, spc = structure(n_cells, seed)
        pair.build_neighbors(cell, pos, spc)
        for form in FORMS:
            print(f"noalloc vs compute: {form}, {spc.numel()} atoms", flush=True)
            ref = run(pair, spc, form, noalloc=False)
            g_ref = torch.empty((2, F, F), dtype=torch.float64, device=DEV)
            pair.beta_grad_device(g_ref)
            out = run(pair, spc, form, noalloc=True)
            g = torch.empty_like(g_ref)
            pair.beta_grad_device(g)
            for a, b, name in zip(out, ref, ("energy", "forces", "virial", "eatom")):
                assert np.all(np.isfinite(a)), (form, name)
                assert np.array_equal(a, b), (form, name, np.abs(a - b).max())
            assert torch.equal(g, g_ref), form


def test_noalloc_refuses_input_beyond_scratch():
    pair = bench_pair()
    # 22^3 fcc cells = 42,592 atoms: the mixed form's workspace (about 7.5 KB
    # per atom) exceeds the 256 MiB scratch
    cell, pos, spc = structure(22, 23)
    pair.build_neighbors(cell, pos, spc)
    n = spc.numel()
    e = torch.zeros(1, dtype=torch.float64, device=DEV)
    f = torch.zeros((n, 3), dtype=torch.float64, device=DEV)
    v = torch.zeros(6, dtype=torch.float64, device=DEV)
    ctx = _capi.context(0)
    rc = _capi.lib().flare_pair_compute_noalloc(
        ctx.handle, pair._h, pair._nlist, _capi.dptr(spc), _capi.FORMS["adjoint_mixed"],
        None, _capi.dptr(e), _capi.dptr(f), _capi.dptr(v), None)
    msg = _capi.lib().flare_last_error().decode()
    print("refusal:", rc, msg, flush=True)
    assert rc == 1 and "too large" in msg
    torch.cuda.synchronize()
    assert float(f.abs().max()) == 0.0 and float(e[0]) == 0.0
    # the scratch is usable again right after the refusal
    cell, pos, spc = structure(2, 24)
    pair.build_neighbors(cell, pos, spc)
    ref = run(pair, spc, "adjoint_mixed", noalloc=False)
    out = run(pair, spc, "adjoint_mixed", noalloc=True)
    for a, b in zip(out, ref):
        assert np.array_equal(a, b)


def test_noalloc_past_1024_columns():
    """F = 1,200 (3 species, n_max 8, l_max 3): the FP32 and mixed forms run
    their contractions on the hand-written kernels and are served; the FP64
    forms would need rocBLAS and are refused before any device work."""
    S, N, L = 3, 8, 3
    F = (S * N) * (S * N + 1) // 2 * (L + 1)
    beta = synth.symmetric_beta(S, F, 1e-3, seed=12)
    packed = np.stack([pack_beta(b) for b in beta])
    pair = PairFLARE.from_coefficients(2, S, N, L, np.full((S, S), 5.0), packed)
    assert pair.n_descriptors == 1200
    for n_cells, seed in ((3, 25), (2, 26)):
        n = 4 * n_cells ** 3
        cell, species, pos = synth.fcc(n_cells, 4.05, 0.08,
                                       [n - 2 * (n // 3), n // 3, n // 3], seed)
        pos = torch.tensor(pos, device=DEV)
        spc = torch.tensor(species, dtype=torch.int32, device=DEV)
        pair.build_neighbors(cell, pos, spc)
        for form in ("adjoint_fp32", "adjoint_mixed"):
            print(f"noalloc vs compute at F = {F}: {form}, {spc.numel()} atoms", flush=True)
            ref = run(pair, spc, form, noalloc=False)
            out = run(pair, spc, form, noalloc=True)
            for a, b, name in zip(out, ref, ("energy", "forces", "virial", "eatom")):
                assert np.all(np.isfinite(a)), (form, name)
                assert np.array_equal(a, b), (form, name, np.abs(a - b).max())
    n = spc.numel()
    f = torch.zeros((n, 3), dtype=torch.float64, device=DEV)
    e = torch.zeros(1, dtype=torch.float64, device=DEV)
    rc = _capi.lib().flare_pair_compute_noalloc(
        _capi.context(0).handle, pair._h, pair._nlist, _capi.dptr(spc),
        _capi.FORMS["adjoint_fp64"], None, _capi.dptr(e), _capi.dptr(f), None, None)
    msg = _capi.lib().flare_last_error().decode()
    print("refusal:", rc, msg, flush=True)
    assert rc == 1 and "F <= 1024" in msg
    torch.cuda.synchronize()
    assert float(f.abs().max()) == 0.0
