"""Seeded synthetic structures and models of the benchmark configurations.

Every generator is deterministic in its seed (numpy PCG64)."""

import numpy as np

FCC_BASIS = np.array([[0.0, 0.0, 0.0], [0.5, 0.5, 0.0], [0.5, 0.0, 0.5],
                      [0.0, 0.5, 0.5]])


def fcc(n_cells, a, sigma, species_counts, seed):
    """Cubic fcc supercell of n_cells^3 conventional cells, lattice constant a,
    Gaussian displacements of std sigma per component, species assigned by a
    seeded shuffle of the exact counts.  Returns cell, species, positions."""
    rng = np.random.default_rng(seed)
    g = np.arange(n_cells)
    cells = np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    frac = (cells[:, None, :] + FCC_BASIS[None, :, :]).reshape(-1, 3)
    pos = frac * a
    n = pos.shape[0]
    pos = pos + rng.normal(0.0, sigma, size=(n, 3))
    counts = list(species_counts)
    if sum(counts) != n:
        raise ValueError(f"species counts {counts} do not sum to {n}")
    species = np.concatenate([np.full(c, s, dtype=np.int32)
                              for s, c in enumerate(counts)])
    species = rng.permutation(species).astype(np.int32)
    cell = np.eye(3) * (n_cells * a)
    return cell, species, pos


def fraction_counts(n, fractions):
    """Exact species counts: floor-rounded shares, remainder to the largest
    fractional parts (0.4/0.3/0.3 of 55,296 -> 22,118 / 16,589 / 16,589)."""
    raw = np.asarray(fractions, dtype=np.float64) * n
    counts = np.floor(raw).astype(int)
    rem = n - counts.sum()
    order = np.argsort(-(raw - counts), kind="stable")
    for k in range(rem):
        counts[order[k]] += 1
    return [int(c) for c in counts]


def symmetric_beta(n_species, F, std, seed):
    """Per-species symmetric beta = (A + A^T)/2 with A ~ N(0, std)."""
    rng = np.random.default_rng(seed)
    out = np.empty((n_species, F, F))
    for s in range(n_species):
        A = rng.normal(0.0, std, size=(F, F))
        out[s] = 0.5 * (A + A.T)
    return out


def config1(seed=1):
    """108-atom single-species fcc, a = 4.05, sigma = 0.05; B2 {1, 8, 3},
    rc 5.0; NormalizedDotProduct(2.0, 2); 50 seeded sparse atoms; alpha ~
    N(0, 1)."""
    cell, species, pos = fcc(3, 4.05, 0.05, [108], seed)
    rng = np.random.default_rng(seed + 1000)
    sparse_atoms = np.sort(rng.choice(108, size=50, replace=False)).tolist()
    alpha = rng.normal(0.0, 1.0, size=50)
    return dict(cell=cell, species=species, positions=pos, cutoff=5.0,
                settings=[1, 8, 3], sigma=2.0, power=2, sparse_atoms=sparse_atoms,
                alpha=alpha)


def config2(seed=2):
    """32,000-atom 2-species fcc, a = 4.05, sigma = 0.08; B2 {2, 8, 4}, rc 5.0;
    per-species symmetric beta 680 x 680 with A ~ N(0, 1e-3) (std)."""
    cell, species, pos = fcc(20, 4.05, 0.08, [16000, 16000], seed)
    beta = symmetric_beta(2, 680, 1e-3, seed + 1000)
    return dict(cell=cell, species=species, positions=pos, cutoff=5.0,
                settings=[2, 8, 4], beta=beta)


def config3(seed=3, n_sparse=500):
    """config2's structure with a GP mean model: 500 seeded sparse atoms,
    alpha ~ N(0, 1), NormalizedDotProduct(2.0, 2)."""
    c = config2(seed)
    rng = np.random.default_rng(seed + 2000)
    sparse_atoms = np.sort(rng.choice(len(c["species"]), size=n_sparse,
                                      replace=False)).tolist()
    c.update(sparse_atoms=sparse_atoms, alpha=rng.normal(0.0, 1.0, n_sparse),
             sigma=2.0, power=2)
    c.pop("beta")
    return c


def config4(seed=4, n_sparse=2000):
    """55,296-atom 3-species fcc, a = 3.6, sigma = 0.1, fractions 0.4/0.3/0.3;
    B2 {3, 8, 3}, rc 4.5; 2,000 seeded sparse atoms."""
    n = 4 * 24 ** 3
    cell, species, pos = fcc(24, 3.6, 0.1, fraction_counts(n, [0.4, 0.3, 0.3]),
                             seed)
    rng = np.random.default_rng(seed + 2000)
    sparse_atoms = np.sort(rng.choice(n, size=n_sparse, replace=False)).tolist()
    return dict(cell=cell, species=species, positions=pos, cutoff=4.5,
                settings=[3, 8, 3], sparse_atoms=sparse_atoms,
                alpha=rng.normal(0.0, 1.0, n_sparse), sigma=2.0, power=2)


def rsa_positions(n, box, d_min, seed):
    """Random sequential addition of n points in a periodic cube (edge box)
    with minimum distance d_min, from the library's deterministic host
    generator (flare_synth_rsa)."""
    import ctypes
    from . import _capi
    pos = np.empty((n, 3), dtype=np.float64)
    _capi.check(_capi.lib().flare_synth_rsa(int(n), float(box), float(d_min),
                                            int(seed), 0,
                                            pos.ctypes.data_as(ctypes.c_void_p)))
    return pos


def config5(seed=5, n=262144, density=0.06, d_min=2.2):
    """Disordered 3-species box: 262,144 atoms at 0.06 atoms/A^3 (cube of
    (n / density)^(1/3) = 163.5 A), random sequential addition with a 2.2 A
    minimum distance, species drawn uniformly; B2 {3, 8, 6}, rc 5.5
    (F = 2,100); per-species symmetric beta with entries of std 1e-4."""
    box = (n / density) ** (1.0 / 3.0)
    pos = rsa_positions(n, box, d_min, seed)
    rng = np.random.default_rng(seed + 1000)
    species = rng.integers(0, 3, n).astype(np.int32)
    beta = symmetric_beta(3, 2100, 1e-4, seed + 2000)
    return dict(cell=np.eye(3) * box, species=species, positions=pos, cutoff=5.5,
                settings=[3, 8, 6], beta=beta)
