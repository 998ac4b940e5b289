"""GPU B2 descriptors and DescriptorValues force derivatives against the
oracle (b2.cpp:76-399)."""
import numpy as np
import pytest

from flare_pp_amd import B2, Structure, synth
from helpers import small_two_species

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b).max() / max(np.abs(b).max(), 1e-300)


def run(orc, cell, species, pos, settings, rc, cutoffs=None, cutoff="quadratic",
        dervs=True):
    S = settings[0]
    b2 = B2("chebyshev", cutoff, [0, rc], [], settings, cutoffs)
    print(f"case n={len(species)} settings={settings}", flush=True)
    s = Structure(cell, species, pos, rc if cutoffs is None else np.max(cutoffs), [b2])
    nl = orc.NeighborList(cell, species, pos, s.cutoff)
    o = orc.B2Result(nl, species, *settings, b2.cutoffs, cutoff=cutoff, dervs=dervs)
    g = o.grouped()
    d = s.descriptors[0]
    off = np.concatenate([[0], np.cumsum(g["n_by_type"])])
    for t in range(S):
        a, b = off[t], off[t + 1]
        assert rel(d.descriptors[t], g["descriptors"][a:b]) < 1e-12
        assert rel(d.descriptor_norms[t], g["norms"][a:b]) < 1e-12
        np.testing.assert_array_equal(d.atom_indices[t], g["atom_indices"][a:b])
    if dervs:
        poff = np.concatenate([[0], np.cumsum(g["neigh_by_type"])])
        for t in range(S):
            a, b = off[t], off[t + 1]
            pa, pb = poff[t], poff[t + 1]
            np.testing.assert_array_equal(d.neighbor_counts[t], g["neighbor_counts"][a:b])
            np.testing.assert_array_equal(d.cumulative_neighbor_counts[t],
                                          g["cum_neighbor_counts"][a:b])
            np.testing.assert_array_equal(d.neighbor_indices[t],
                                          g["neighbor_indices"][pa:pb])
            np.testing.assert_array_equal(d.neighbor_coordinates[t],
                                          g["neighbor_coords"][pa:pb])
            assert rel(d.descriptor_force_dervs[t], g["force_dervs"][3 * pa:3 * pb]) < 1e-11
            assert rel(d.descriptor_force_dots[t], g["force_dots"][3 * pa:3 * pb]) < 1e-11
    return s, o


def test_config1(orc, cfg1):
    c = cfg1
    run(orc, c["cell"], c["species"], c["positions"], c["settings"], 5.0)


def test_two_species_multicut_cosine(orc):
    cell, species, pos = small_two_species(seed=3)
    run(orc, cell, species, pos, [2, 5, 4], 5.0,
        cutoffs=np.array([[5.0, 4.4], [4.0, 4.7]]), cutoff="cosine")


def test_config2_descriptors(orc):
    c = synth.config2()
    run(orc, c["cell"], c["species"], c["positions"], c["settings"], 5.0, dervs=False)


def test_high_lmax(orc):
    cell, species, pos = small_two_species(seed=4)
    run(orc, cell, species, pos, [2, 3, 8], 5.0)
