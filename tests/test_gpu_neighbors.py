"""GPU neighbor list and species grouping against the oracle: the integer
arrays bit-exact, the relative vectors and wrapped positions bitwise
(structure.cpp:27-43, 52-133)."""
import numpy as np
import pytest

from flare_pp_amd import Structure, synth

pytestmark = pytest.mark.gpu


def compare(orc, cell, species, pos, cutoff):
    print(f"case n={len(species)} cutoff={cutoff}", flush=True)
    s = Structure(cell, species, pos, cutoff)
    o = orc.NeighborList(cell, species, pos, cutoff, "cells")
    assert s.n_neighbors == o.n_pairs
    np.testing.assert_array_equal(s.neighbor_count, o.count)
    np.testing.assert_array_equal(s.cumulative_neighbor_count, o.cum)
    np.testing.assert_array_equal(s.structure_indices, o.idx)
    np.testing.assert_array_equal(s.neighbor_species, o.nspec)
    np.testing.assert_array_equal(s.relative_positions, o.rel)
    np.testing.assert_array_equal(s.wrapped_positions, o.wrapped)
    assert s.sweep == o.sweep
    assert s.volume == orc.volume(cell)
    return s, o


def test_config1(orc, cfg1):
    c = cfg1
    compare(orc, c["cell"], c["species"], c["positions"], 5.0)


@pytest.mark.parametrize("cutoff", [3.0, 5.5])
def test_triclinic_small_cell(orc, cutoff):
    rng = np.random.default_rng(11)
    cell = np.array([[3.1, 0.0, 0.0], [0.9, 2.7, 0.0], [0.4, -0.6, 4.2]])
    pos = rng.random((7, 3)) @ cell + rng.normal(0, 0.3, (7, 3))
    species = rng.integers(0, 2, 7)
    compare(orc, cell, species, pos, cutoff)


def test_isolated_and_empty(orc):
    cell = np.eye(3) * 30.0
    pos = np.array([[1.0, 1.0, 1.0], [15.0, 15.0, 15.0], [17.0, 15.5, 15.0]])
    s, _ = compare(orc, cell, [0, 1, 0], pos, 5.0)
    assert s.neighbor_count[0] == 0


def test_config2_full(orc):
    c = synth.config2()
    s, o = compare(orc, c["cell"], c["species"], c["positions"], 5.0)
    # reverse pairs: (i -> j, s) and (j -> i, -s) have opposite vectors
    from flare_pp_amd.structure import d2h
    rev = d2h(s._view.d_reverse_pair, s.n_neighbors, np.int32)
    assert rev.min() >= 0
    np.testing.assert_array_equal(o.rel[rev, 1:], -o.rel[:, 1:])
    np.testing.assert_array_equal(rev[rev], np.arange(len(rev)))


def test_above_block_scan_limit(orc):
    """275,684 atoms (41^3 fcc cells, a = 3.6 A, rc 4.5 A): above the
    262,144 atoms whose CSR offsets come from the enumeration's per-block
    sums, so the device-wide count scan runs; bitwise against the oracle's
    cell-list build, reverse pairs an involution with opposite vectors."""
    n = 4 * 41 ** 3
    cell, species, pos = synth.fcc(41, 3.6, 0.05, [n // 2, n - n // 2], 57)
    s, o = compare(orc, cell, species, pos, 4.5)
    from flare_pp_amd.structure import d2h
    rev = d2h(s._view.d_reverse_pair, s.n_neighbors, np.int32)
    assert rev.min() >= 0
    np.testing.assert_array_equal(o.rel[rev, 1:], -o.rel[:, 1:])
    np.testing.assert_array_equal(rev[rev], np.arange(len(rev)))


def test_species_grouping(orc):
    c = synth.config4()
    s = Structure(c["cell"], c["species"], c["positions"], 4.5)
    rows = s._atom_of_row()
    want = np.concatenate([np.flatnonzero(np.asarray(c["species"]) == k)
                           for k in range(3)])
    np.testing.assert_array_equal(rows, want)
    assert list(s.species_count[:3]) == [22118, 16589, 16589]


def test_dense_two_pass_path(orc):
    # about 190 neighbors per atom: more than the fused sort keeps in LDS
    # (128), so the two-pass enumerate / rank-sort / reverse path runs
    cell, species, pos = synth.fcc(4, 4.05, 0.05, [256], 5)
    s, o = compare(orc, cell, species, pos, 9.0)
    assert s.neighbor_count.max() > 128
    from flare_pp_amd.structure import d2h
    rev = d2h(s._view.d_reverse_pair, s.n_neighbors, np.int32)
    np.testing.assert_array_equal(o.rel[rev, 1:], -o.rel[:, 1:])


def test_fused_path_reverse_pairs(orc, cfg1):
    c = cfg1
    s, o = compare(orc, c["cell"], c["species"], c["positions"], 5.0)
    assert s.neighbor_count.max() <= 128
    from flare_pp_amd.structure import d2h
    rev = d2h(s._view.d_reverse_pair, s.n_neighbors, np.int32)
    assert rev.min() >= 0
    np.testing.assert_array_equal(o.rel[rev, 1:], -o.rel[:, 1:])
    np.testing.assert_array_equal(rev[rev], np.arange(len(rev)))


def random_box(seed, rho, L, cell=None):
    rng = np.random.default_rng(seed)
    cell = np.eye(3) * L if cell is None else np.asarray(cell, dtype=np.float64)
    n = int(round(rho * abs(np.linalg.det(cell))))
    pos = rng.random((n, 3)) @ cell
    species = rng.integers(0, 2, n)
    return cell, species, pos


def test_bin_staging_limit(orc):
    # six bins of 6 A per axis, ~0.082 atoms/A^3: the 27 bins around a home
    # bin hold ~478 candidates on average, so home bins fall on both sides of
    # the 480 staged in LDS (above it, the per-atom pass from global memory
    # runs); ~74 neighbors per atom, below the 128 the rows keep
    cell, species, pos = random_box(21, 0.082, 36.0)
    s, o = compare(orc, cell, species, pos, 6.0)
    assert s.neighbor_count.max() <= 128
    from flare_pp_amd.structure import d2h
    rev = d2h(s._view.d_reverse_pair, s.n_neighbors, np.int32)
    np.testing.assert_array_equal(rev[rev], np.arange(len(rev)))
    np.testing.assert_array_equal(o.rel[rev, 1:], -o.rel[:, 1:])


def test_bins_over_capacity(orc):
    # ~138 neighbors per atom with at least three bins per axis: the staged
    # kernel's fallback flags the atoms above 128 and the two-pass path
    # rebuilds the list
    cell, species, pos = random_box(22, 0.12, 39.0)
    s, o = compare(orc, cell, species, pos, 6.5)
    assert s.neighbor_count.max() > 128


def test_triclinic_bins(orc):
    # sheared cell with three or more bins along every axis
    cell = np.array([[24.0, 0.0, 0.0], [5.0, 22.0, 0.0], [-3.0, 4.0, 23.0]])
    cell, species, pos = random_box(23, 0.05, None, cell)
    compare(orc, cell, species, pos, 5.0)


def test_prefilter_at_the_cutoff(orc):
    # pairs at distances within 1e-15 relative of the cutoff, on both sides,
    # in a cell whose coordinates reach ~1,000 A (FP32 positions there are
    # rounded by ~6e-5 A): the FP32 prefilter must keep every pair the exact
    # FP64 test keeps, so the list equals the oracle's bit for bit
    rng = np.random.default_rng(31)
    L, rc = 1000.0, 5.0
    cell = np.diag([L, 40.0, 37.0])
    pos = []
    for t in range(150):
        c = np.array([rng.uniform(0, L), rng.uniform(0, 40.0), rng.uniform(0, 37.0)])
        u = rng.normal(size=3)
        u /= np.linalg.norm(u)
        for f in (1 - 2e-16, 1 - 1e-15, 1.0, 1 + 1e-15, 1 - 1e-7, 1 + 1e-7):
            pos.append(c)
            pos.append(c + rc * f * u)
    pos = np.array(pos)
    species = rng.integers(0, 2, len(pos))
    s, o = compare(orc, cell, species, pos, rc)
    assert s.neighbor_count.sum() > 0


@pytest.mark.parametrize("shape", [(3, 3, 3), (3, 7, 4)])
def test_three_bins_per_axis(orc, shape):
    # exactly three bins along an axis: the bins on both sides of a home bin
    # are the other two, reached with opposite image shifts
    rc = 5.0
    cell = np.diag([3 * rc + 0.4 + b for b in (0.0, 0.1, 0.2)]) * np.array(shape) / 3.0
    cell, species, pos = random_box(40 + shape[1], 0.04, None, cell)
    compare(orc, cell, species, pos, rc)


def test_crowded_bin_regroups(orc):
    """An fcc box (4 bins per axis) with 40 extra atoms packed into one
    bin: more atoms than a fixed-capacity bin holds (32), so the build
    groups the atoms again in the scanned layout and enumerates again.
    Bitwise against the oracle."""
    cell, species, pos = synth.fcc(5, 4.05, 0.05, [250, 250], 71)
    rng = np.random.default_rng(72)
    extra = []
    while len(extra) < 40:
        p = rng.uniform(0.6, 4.4, 3)
        if all(np.linalg.norm(p - q) > 0.9 for q in extra):
            extra.append(p)
    pos = np.vstack([pos, np.array(extra)])
    species = np.concatenate([species, rng.integers(0, 2, 40)])
    frac = np.linalg.solve(cell.T, pos.T).T % 1.0
    bins = np.floor(frac * 4).astype(int)
    occupancy = np.bincount((bins[:, 0] * 4 + bins[:, 1]) * 4 + bins[:, 2], minlength=64)
    assert occupancy.max() > 32
    s, o = compare(orc, cell, species, pos, 5.0)
    from flare_pp_amd.structure import d2h
    rev = d2h(s._view.d_reverse_pair, s.n_neighbors, np.int32)
    assert rev.min() >= 0
    np.testing.assert_array_equal(o.rel[rev, 1:], -o.rel[:, 1:])
