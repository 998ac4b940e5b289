"""The context's execution modes give the same results as the default one:
work enqueued on a caller's stream (flare_ctx_set_stream), calls that return
once enqueued (flare_ctx_set_blocking 0, then flare_ctx_synchronize), stage
timing (flare_ctx_profile), and repeated calls.  Forces are assembled by a
gather over reverse pairs with no floating-point atomics, so every mode and
every repetition gives the same bits."""
import numpy as np
import pytest
import torch

from flare_pp_amd import PairFLARE, _capi, pack_beta, synth
from helpers import small_two_species

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
FORMS = ("plain", "adjoint_fp64", "adjoint_fp32", "adjoint_mixed")


@pytest.fixture(scope="module")
def system():
    F = 16 * 17 // 2 * 4
    beta = synth.symmetric_beta(2, F, 1e-3, 3)
    pair = PairFLARE.from_coefficients(2, 2, 8, 3, np.full((2, 2), 5.0),
                                       np.stack([pack_beta(b) for b in beta]))
    cell, species, pos = small_two_species(seed=21, n_cells=3)
    ref = {f: pair.compute(cell, species, pos, form=f) for f in FORMS}
    return pair, cell, species, pos, ref


def same(a, b):
    return a[0] == b[0] and np.array_equal(a[1], b[1]) and np.array_equal(a[2], b[2])


@pytest.mark.parametrize("form", FORMS)
def test_repeated_calls_give_the_same_bits(system, form):
    pair, cell, species, pos, ref = system
    print(f"repeat, form {form}", flush=True)
    for _ in range(3):
        assert same(pair.compute(cell, species, pos, form=form), ref[form])


@pytest.mark.parametrize("form", ["adjoint_fp64", "adjoint_mixed"])
def test_caller_stream(system, form):
    # torch's current stream is a side stream: the neighbor build, the
    # forward and force kernels and the GEMM all go to it
    pair, cell, species, pos, ref = system
    print(f"side stream, form {form}", flush=True)
    s = torch.cuda.Stream()
    with torch.cuda.stream(s):
        got = pair.compute(cell, species, pos, form=form)
    torch.cuda.synchronize()
    assert same(got, ref[form])


@pytest.mark.parametrize("form", FORMS)
def test_non_blocking_calls(system, form):
    pair, cell, species, pos, ref = system
    print(f"non-blocking, form {form}", flush=True)
    ctx = _capi.context(0)
    n = len(species)
    d_pos = torch.as_tensor(pos, dtype=torch.float64, device=DEV)
    d_spc = torch.as_tensor(np.asarray(species, dtype=np.int32), device=DEV)
    e = torch.empty(1, dtype=torch.float64, device=DEV)
    f = torch.empty((n, 3), dtype=torch.float64, device=DEV)
    v = torch.empty(6, dtype=torch.float64, device=DEV)
    ctx.set_blocking(False)
    try:
        for _ in range(2):
            pair.build_neighbors(cell, d_pos, d_spc)
            pair.compute_device(d_spc, form=form, energy=e, forces=f, virial=v)
        ctx.synchronize()
    finally:
        ctx.set_blocking(True)
    assert same((e.item(), f.cpu().numpy(), v.cpu().numpy()), ref[form])


def test_stage_profile(system):
    pair, cell, species, pos, ref = system
    print("stage profile", flush=True)
    ctx = _capi.context(0)
    ctx.profile(True)
    try:
        got = pair.compute(cell, species, pos, form="adjoint_mixed")
        st = ctx.stage_times()
    finally:
        ctx.profile(False)
    assert same(got, ref["adjoint_mixed"])
    for k in ("nlist", "forward", "gemm", "forces_adjoint", "gather_reduce"):
        assert k in st and st[k][1] == 1 and st[k][0] > 0, (k, st)
    assert ctx.stage_times() == {}
