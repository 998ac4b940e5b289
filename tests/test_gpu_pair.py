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
        e, f, v = pair.compute(cell, species, pos, form=form)
        print(f"form {form} n={len(species)}: force max rel {nrel(f, f0):.2e}, "
              f"RMS rel {rms_rel(f, f0):.2e}, energy rel "
              f"{abs(e - e0) / np.abs(ea0).sum():.2e}", flush=True)
        assert nrel(f, f0) < mixed_tol, form
        assert abs(e - e0) < mixed_tol * np.abs(ea0).sum(), form
    return e0, f0, v0


def rms_rel(f, f0):
    return float(np.sqrt(np.mean((f - f0) ** 2)) / np.sqrt(np.mean(f0 ** 2)))


# Error bounds of the low-precision forms against FP64, relative to the
# reference's scale: FP32 rows and contractions keep ~1e-7 relative per
# value, the mixed form's BF16 hi + lo products ~2^-16 per operand.
LOWP_BOUNDS = {
    # form: (force RMS rel, force max rel, energy rel to sum |E_i|, virial rel)
    "adjoint_fp32": (1e-5, 1e-4, 1e-5, 1e-4),
    "adjoint_mixed": (5e-5, 5e-4, 1e-5, 1e-4),
}


def check_low_precision(pair, cell, species, pos, e0, f0, v0, ea0):
    for form, (b_rms, b_max, b_e, b_v) in LOWP_BOUNDS.items():
        e, f, v = pair.compute(cell, species, pos, form=form)
        er = (rms_rel(f, f0), nrel(f, f0), abs(e - e0) / np.abs(ea0).sum(), nrel(v, v0))
        print(f"{form} n={len(species)}: force RMS rel {er[0]:.2e} max rel {er[1]:.2e} "
              f"energy rel {er[2]:.2e} virial rel {er[3]:.2e} (force RMS "
              f"{np.sqrt(np.mean(f0 ** 2)) * 1e3:.3e} meV/A)", flush=True)
        assert er[0] < b_rms and er[1] < b_max and er[2] < b_e and er[3] < b_v, form


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


def test_wide_gemm_ragged_tiles(orc):
    """The mixed form's wide-row GEMM (features > 256) with a feature count
    that is neither a multiple of the 352-column tile nor of 4 (B2 {2, 7, 4},
    F = 525: a full and a 173-column tile, scalar Z stores) and species of
    300 and 200 atoms (a partial 256-row tile): every form against the
    oracle, the low-precision forms to the stated bounds."""
    cell, species, pos = synth.fcc(5, 4.05, 0.1, [300, 200], 57)
    pair = make_pair(2, 7, 4, seed=58)
    assert pair.n_descriptors == 525
    e0, f0, v0 = check_forms(orc, pair, cell, species, pos)
    ea0 = oracle_pair(orc, pair, cell, species, pos)[3]
    check_low_precision(pair, cell, species, pos, e0, f0, v0, ea0)


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
    cell, species, pos = c["cell"], c["species"], c["positions"]
    e0, f0, v0, ea0 = oracle_pair(orc, pair, cell, species, pos)
    for form in ("plain", "adjoint_fp64"):
        e, f, v, ea = pair.compute(cell, species, pos, form=form, want_eatom=True)
        assert abs(e - e0) <= 1e-10 * np.abs(ea0).sum(), form
        assert nrel(f, f0) < 1e-10 and nrel(v, v0) < 1e-10 and nrel(ea, ea0) < 1e-10, form
    # the synthetic beta gives RMS forces of ~1 meV/A, so the bounds are
    # relative: an absolute 1 meV/A gate would pass zero forces
    check_low_precision(pair, cell, species, pos, e0, f0, v0, ea0)


def test_config5_settings_small_box(orc):
    """Config 5's settings (3 species, B2 {3, 8, 6}, rc 5.5, F = 2,100) on a
    400-atom random-sequential-addition box at its density (0.06 atoms/A^3,
    2.2 A minimum distance): the compile-time (n_max 8, l_max 6) forward and
    adjoint force kernels against the oracle in every form."""
    n = 400
    box = (n / 0.06) ** (1.0 / 3.0)
    pos = synth.rsa_positions(n, box, 2.2, seed=51)
    species = np.random.default_rng(52).integers(0, 3, n).astype(np.int32)
    cell = np.eye(3) * box
    beta = synth.symmetric_beta(3, 2100, 1e-4, 53)
    pair = PairFLARE.from_coefficients(2, 3, 8, 6, np.full((3, 3), 5.5),
                                       np.stack([pack_beta(b) for b in beta]))
    e0, f0, v0, ea0 = oracle_pair(orc, pair, cell, species, pos)
    for form in ("plain", "adjoint_fp64"):
        e, f, v, ea = pair.compute(cell, species, pos, form=form, want_eatom=True)
        print(f"{form}: force rel {nrel(f, f0):.2e} energy rel "
              f"{abs(e - e0) / np.abs(ea0).sum():.2e}", flush=True)
        assert abs(e - e0) <= 1e-10 * np.abs(ea0).sum(), form
        assert nrel(f, f0) < 1e-10 and nrel(v, v0) < 1e-10 and nrel(ea, ea0) < 1e-10, form
    check_low_precision(pair, cell, species, pos, e0, f0, v0, ea0)


@pytest.mark.parametrize("S", [1, 2])
def test_lmax6_fewer_channels(orc, S):
    """The l_max 6 force kernel that writes u over A with 8 and 16 channels
    (B2 {1, 8, 6}, F = 252; B2 {2, 8, 6}, F = 952) on a 300-atom
    random-sequential-addition box: every form against the oracle."""
    n = 300
    box = (n / 0.06) ** (1.0 / 3.0)
    pos = synth.rsa_positions(n, box, 2.2, seed=60 + S)
    species = np.random.default_rng(62 + S).integers(0, S, n).astype(np.int32)
    cell = np.eye(3) * box
    F = (S * 8) * (S * 8 + 1) // 2 * 7
    beta = synth.symmetric_beta(S, F, 1e-4, 64 + S)
    pair = PairFLARE.from_coefficients(2, S, 8, 6, np.full((S, S), 5.5),
                                       np.stack([pack_beta(b) for b in beta]))
    e0, f0, v0, ea0 = oracle_pair(orc, pair, cell, species, pos)
    for form in ("plain", "adjoint_fp64"):
        e, f, v, ea = pair.compute(cell, species, pos, form=form, want_eatom=True)
        assert abs(e - e0) <= 1e-10 * np.abs(ea0).sum(), form
        assert nrel(f, f0) < 1e-10 and nrel(v, v0) < 1e-10 and nrel(ea, ea0) < 1e-10, form
    check_low_precision(pair, cell, species, pos, e0, f0, v0, ea0)


def test_dense_environments_mixed(orc):
    """About 190 neighbors per atom (rc 9 A in fcc at 4.05 A; the two-pass
    neighbor build): the benchmark's B2 {2, 8, 4} kernels in every form
    against the oracle, the low-precision forms to the stated bounds."""
    cell, species, pos = synth.fcc(4, 4.05, 0.05, [128, 128], 55)
    S, N, L = 2, 8, 4
    F = (S * N) * (S * N + 1) // 2 * (L + 1)
    beta = synth.symmetric_beta(S, F, 1e-3, 56)
    pair = PairFLARE.from_coefficients(2, S, N, L, np.full((S, S), 9.0),
                                       np.stack([pack_beta(b) for b in beta]))
    e0, f0, v0, ea0 = oracle_pair(orc, pair, cell, species, pos)
    for form in ("plain", "adjoint_fp64"):
        e, f, v, ea = pair.compute(cell, species, pos, form=form, want_eatom=True)
        assert abs(e - e0) <= 1e-10 * np.abs(ea0).sum(), form
        assert nrel(f, f0) < 1e-10 and nrel(v, v0) < 1e-10, form
    check_low_precision(pair, cell, species, pos, e0, f0, v0, ea0)


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
    # with it norm-wise to 1e-10 and the FP32 and mixed forms stay within the
    # relative bounds.  The same settings are pinned to the oracle on the
    # 400-atom box of test_config5_settings_small_box.
    c = synth.config5()
    packed = np.stack([pack_beta(b) for b in c["beta"]])
    pair = PairFLARE.from_coefficients(2, 3, 8, 6, np.full((3, 3), 5.5), packed)
    cell, species, pos = c["cell"], c["species"], c["positions"]
    n = len(species)
    print(f"config5 n={n}", flush=True)
    e0, f0, v0, ea0 = pair.compute(cell, species, pos, form="plain", want_eatom=True)
    assert np.abs(f0.sum(axis=0)).max() < 1e-9 * np.abs(f0).sum()
    e1, f1, v1 = pair.compute(cell, species, pos, form="adjoint_fp64")
    assert abs(e1 - e0) < 1e-10 * abs(e0) and nrel(f1, f0) < 1e-10
    assert nrel(v1, v0) < 1e-10
    for form, (b_rms, b_max, b_e, b_v) in LOWP_BOUNDS.items():
        e2, f2, v2 = pair.compute(cell, species, pos, form=form)
        d = (f2 - f0) * 1e3
        print(f"{form}: energy {abs(e2 - e0) / n * 1e3:.3e} meV/atom, force RMS "
              f"{np.sqrt(np.mean(d ** 2)):.3e} max {np.abs(d).max():.3e} meV/A, "
              f"RMS rel {rms_rel(f2, f0):.2e}", flush=True)
        # relative bounds: config 5's RMS force is ~0.2 meV/A
        assert rms_rel(f2, f0) < b_rms and nrel(f2, f0) < b_max, form
        assert abs(e2 - e0) < b_e * np.abs(ea0).sum() and nrel(v2, v0) < b_v, form


def test_rebuild_sequence_one_list(orc):
    """One neighbor list rebuilt for a run of different structures: larger
    (its buffers grow), smaller (counters left zeroed by the previous
    build's compaction, no clearing launch), a thin triclinic cell (the
    per-atom enumeration), a dense box with more than 128 neighbors per atom
    (the two-pass rebuild), then the first structure again.  Every build in
    the adjoint FP64 form against the oracle (1e-10), the mixed form on the
    last."""
    pair = make_pair(2, 4, 3, seed=61)
    rng = np.random.default_rng(62)
    thin_cell = np.array([[3.1, 0.0, 0.0], [0.9, 2.7, 0.0], [0.4, -0.6, 4.2]])
    thin_pos = rng.random((7, 3)) @ thin_cell + rng.normal(0, 0.3, (7, 3))
    cases = [
        ("fcc 2^3", *small_two_species(seed=63)),
        ("fcc 5^3", *synth.fcc(5, 4.05, 0.08, [250, 250], 64)),
        ("fcc 3^3", *synth.fcc(3, 4.05, 0.08, [54, 54], 65)),
        ("thin triclinic", thin_cell, rng.integers(0, 2, 7), thin_pos),
        ("dense a = 2.5", *synth.fcc(4, 2.5, 0.03, [128, 128], 66)),
        ("fcc 2^3 again", *small_two_species(seed=63)),
    ]
    for name, cell, species, pos in cases:
        e0, f0, v0, ea0 = oracle_pair(orc, pair, cell, species, pos)
        e, f, v, ea = pair.compute(cell, species, pos, form="adjoint_fp64", want_eatom=True)
        print(f"{name}: n={len(species)} force rel {nrel(f, f0):.2e}", flush=True)
        assert abs(e - e0) <= 1e-10 * np.abs(ea0).sum(), name
        assert nrel(f, f0) < 1e-10 and nrel(v, v0) < 1e-10 and nrel(ea, ea0) < 1e-10, name
    check_low_precision(pair, cell, species, pos, e0, f0, v0, ea0)
