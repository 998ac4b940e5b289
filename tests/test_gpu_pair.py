"""GPU mapped pair style against the oracle restatement of pair_flare
(pair_flare.cpp:68-184): plain FP64 and adjoint FP64 to 1e-10 norm-wise,
FP32 / mixed forms within their stated error."""
import os

import numpy as np
import pytest

from conftest import GOLDEN
from flare_pp_amd import PairFLARE, pack_beta, synth
from helpers import small_two_species

pytestmark = pytest.mark.gpu


def nrel(a, b):
    return np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300)


def oracle_pair(orc, pair, cell, species, pos):
    nl = orc.NeighborList(cell, species, pos, pair.cutoff)
    return orc.pair_compute(nl, species, pair.n_species, pair.n_max, pair.l_max,
                            pair.power, pair.cutoff_matrix, pair.beta,
                            cutoff=pair.cutoff_function, packed=False,
                            want_eatom=True)


def check_forms(orc, pair, cell, species, pos, mixed_tol=1e-3):
    e0, f0, v0, ea0 = oracle_pair(orc, pair, cell, species, pos)
    for form in ("plain", "adjoint_fp64"):
        print(f"form {form} n={len(species)}", flush=True)
        e, f, v, ea = pair.compute(cell, species, pos, form=form, want_eatom=True)
        assert abs(e - e0) <= 1e-10 * max(abs(e0), np.abs(ea0).max()), form
        assert nrel(f, f0) < 1e-10, form
        assert nrel(v, v0) < 1e-10, form
        assert nrel(ea, ea0) < 1e-10, form
    for form in ("adjoint_fp32", "adjoint_mixed"):
        print(f"form {form} n={len(species)}", flush=True)
        e, f, v = pair.compute(cell, species, pos, form=form)
        assert nrel(f, f0) < mixed_tol, form
        assert abs(e - e0) < mixed_tol * np.abs(ea0).sum(), form
    return e0, f0, v0


def make_pair(S, N, L, seed, power=2, std=1e-3, cutoffs=None, cutoff="quadratic"):
    F = (S * N) * (S * N + 1) // 2 * (L + 1)
    cut = np.full((S, S), 5.0) if cutoffs is None else cutoffs
    if power == 2:
        beta = synth.symmetric_beta(S, F, std, seed)
        packed = np.stack([pack_beta(b) for b in beta])
    else:
        packed = np.random.default_rng(seed).normal(0, std, (S, F))
    return PairFLARE.from_coefficients(power, S, N, L, cut, packed,
                                       cutoff_function=cutoff)


@pytest.mark.parametrize("power", [1, 2])
def test_two_species_forms(orc, power):
    cell, species, pos = small_two_species(seed=7)
    check_forms(orc, make_pair(2, 4, 3, seed=2, power=power), cell, species, pos)


def test_multicut_cosine(orc):
    cell, species, pos = small_two_species(seed=8)
    pair = make_pair(2, 5, 4, seed=3, cutoffs=np.array([[5.0, 4.4], [4.0, 4.7]]),
                     cutoff="cosine")
    check_forms(orc, pair, cell, species, pos)


def test_legacy_fixture_si(orc):
    # Si.txt: legacy layout (no power line), 1 species, N 8, L 4, rc 5.0
    pair = PairFLARE(os.path.join(GOLDEN, "Si.txt"))
    assert pair.legacy and pair.power == 2 and pair.n_descriptors == 180
    # diamond silicon, 2 x 2 x 2 conventional cells, displaced
    a = 5.431
    basis = np.array([[0, 0, 0], [0, .5, .5], [.5, 0, .5], [.5, .5, 0],
                      [.25, .25, .25], [.25, .75, .75], [.75, .25, .75],
                      [.75, .75, .25]])
    g = np.arange(2)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    pos = ((cells[:, None] + basis[None]) * a).reshape(-1, 3)
    pos += np.random.default_rng(0).normal(0, 0.05, pos.shape)
    cell = np.eye(3) * 2 * a
    species = np.zeros(len(pos), dtype=np.int32)
    check_forms(orc, pair, cell, species, pos, mixed_tol=5e-3)


def test_legacy_fixture_two_species_cosine(orc):
    pair = PairFLARE(os.path.join(GOLDEN, "beta.txt"))
    assert pair.cutoff_function == "cosine" and pair.n_species == 2
    cell, species, pos = small_two_species(seed=9)
    check_forms(orc, pair, cell, species, pos, mixed_tol=5e-3)


def test_empty_environment(orc):
    cell = np.eye(3) * 30.0
    pos = np.array([[1.0, 1.0, 1.0], [15.0, 15.0, 15.0], [17.0, 15.5, 15.0]])
    pair = make_pair(2, 3, 2, seed=4)
    e, f, v, ea = pair.compute(cell, [0, 1, 0], pos, want_eatom=True)
    assert ea[0] == 0 and np.all(f[0] == 0) and np.isfinite(e)
    e0, f0, v0, ea0 = oracle_pair(orc, pair, cell, [0, 1, 0], pos)
    assert nrel(f, f0) < 1e-10


def test_config2_full(orc):
    c = synth.config2()
    packed = np.stack([pack_beta(b) for b in c["beta"]])
    pair = PairFLARE.from_coefficients(2, 2, 8, 4, np.full((2, 2), 5.0), packed)
    e0, f0, v0 = check_forms(orc, pair, c["cell"], c["species"], c["positions"])
    for form in ("adjoint_fp32", "adjoint_mixed"):
        e, f, v = pair.compute(c["cell"], c["species"], c["positions"], form=form)
        rms = np.sqrt(np.mean((f - f0) ** 2))
        print(f"{form}: dE/atom {abs(e - e0) / len(f) * 1e3:.3e} meV, "
              f"force RMS {rms * 1e3:.3e} meV/A max {np.abs(f - f0).max() * 1e3:.3e}",
              flush=True)
        assert rms < 1e-3  # 1 meV/A


def test_plain_form_in_atom_ranges(orc):
    # a derivative budget of ~40 pairs forces the plain form through many
    # atom ranges (and single-atom ranges); results must not change
    from flare_pp_amd import _capi
    cell, species, pos = small_two_species(seed=9)
    pair = make_pair(2, 4, 3, seed=4)
    e0, f0, v0 = pair.compute(cell, species, pos, form="plain")[:3]
    ctx = _capi.context(0)
    per_pair = 3 * (4 * 16) * 8
    try:
        for budget in (40 * per_pair, 1):
            ctx.set_plain_chunk(budget)
            e, f, v = pair.compute(cell, species, pos, form="plain")[:3]
            assert e == e0 and np.array_equal(f, f0) and np.array_equal(v, v0)
    finally:
        ctx.set_plain_chunk(8 << 30)
    e1, f1, v1, _ = oracle_pair(orc, pair, cell, species, pos)
    assert nrel(f0, f1) < 1e-10


def test_config5_scaleup_forms():
    # 262,144-atom 3-species RSA liquid, l_max 6 (F = 2,100): the plain form
    # runs through atom ranges (its dA/dr is ~100 GB); the FP64 adjoint agrees
    # with it norm-wise to 1e-10 and the mixed form stays within 1 meV/A RMS.
    # The plain form is pinned to the oracle on the smaller cases above.
    c = synth.config5()
    packed = np.stack([pack_beta(b) for b in c["beta"]])
    pair = PairFLARE.from_coefficients(2, 3, 8, 6, np.full((3, 3), 5.5), packed)
    cell, species, pos = c["cell"], c["species"], c["positions"]
    n = len(species)
    print(f"config5 n={n}", flush=True)
    e0, f0, v0 = pair.compute(cell, species, pos, form="plain")
    assert np.abs(f0.sum(axis=0)).max() < 1e-9 * np.abs(f0).sum()
    e1, f1, v1 = pair.compute(cell, species, pos, form="adjoint_fp64")
    assert abs(e1 - e0) < 1e-10 * abs(e0) and nrel(f1, f0) < 1e-10
    assert nrel(v1, v0) < 1e-10
    e2, f2, v2 = pair.compute(cell, species, pos, form="adjoint_mixed")
    d = (f2 - f0) * 1e3
    rms = float(np.sqrt(np.mean(d ** 2)))
    print(f"mixed: energy {abs(e2 - e0) / n * 1e3:.3e} meV/atom, force RMS "
          f"{rms:.3e} max {np.abs(d).max():.3e} meV/A", flush=True)
    assert rms < 1.0
