"""Small seeded systems shared by the tests."""
import numpy as np

from flare_pp_amd import synth


def small_two_species(seed=0, n_cells=2, a=4.05, sigma=0.1):
    cell, species, pos = synth.fcc(n_cells, a, sigma,
                                   [2 * n_cells ** 3, 2 * n_cells ** 3], seed)
    return cell, species, pos


def isolated_cluster(seed=0, n=6, box=30.0):
    """A few atoms in a large box: every neighbor is a distinct atom."""
    rng = np.random.default_rng(seed)
    pos = []
    while len(pos) < n:
        p = rng.uniform(-2.2, 2.2, 3)
        if all(np.linalg.norm(p - q) > 1.6 for q in pos):
            pos.append(p)
    pos = np.array(pos) + box / 2
    species = rng.integers(0, 2, n).astype(np.int32)
    return np.eye(3) * box, species, pos


def strain(cell, pos, a, b, eps):
    e = np.zeros((3, 3))
    e[a, b] = eps
    m = np.eye(3) + e
    return cell @ m.T, pos @ m.T


def grouped_sparse(groups, atoms_by_type):
    """Sparse set = selected clusters of a grouped descriptor dict."""
    rows, counts = [], []
    off = np.concatenate([[0], np.cumsum(groups["n_by_type"])])
    for s, sel in enumerate(atoms_by_type):
        rows.extend(off[s] + np.asarray(sel, dtype=int))
        counts.append(len(sel))
    rows = np.asarray(rows, dtype=int)
    return groups["descriptors"][rows], groups["norms"][rows], counts
