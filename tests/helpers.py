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


def lammps_ghost_list(cell, species, positions, ghost_cut, list_cut, seed=0,
                      shuffle=True):
    """A stand-in for LAMMPS' atom/comm/neighbor setup of one rank owning the
    whole periodic box: wrapped local atoms first, then ghost images within
    ghost_cut of the box (fractional test against the perpendicular widths),
    1-based types, and a full neighbor list of every local atom (all atoms
    within list_cut, in a seeded random order, as LAMMPS' binned order is
    arbitrary).  Returns a dict with x (n_all x 3), type, owner (n_all -> local
    atom), ilist, numneigh (per local atom), offsets/neigh (CSR) and the
    padded 2-D array nbr2d (n_local x maxneigh, row = atom)."""
    from scipy.spatial import cKDTree
    from oracle import pyoracle

    cell = np.asarray(cell, dtype=np.float64)
    n = len(species)
    w = pyoracle.wrap_positions(cell, np.asarray(positions, dtype=np.float64))
    inv = np.linalg.inv(cell)
    frac = w @ inv
    # perpendicular widths: volume / |a_j x a_k|
    V = abs(np.linalg.det(cell))
    widths = np.array([V / np.linalg.norm(np.cross(cell[(k + 1) % 3], cell[(k + 2) % 3]))
                       for k in range(3)])
    fmargin = ghost_cut / widths
    reps = np.ceil(fmargin).astype(int)
    xs, owner = [w], [np.arange(n)]
    for s1 in range(-reps[0], reps[0] + 1):
        for s2 in range(-reps[1], reps[1] + 1):
            for s3 in range(-reps[2], reps[2] + 1):
                if s1 == 0 and s2 == 0 and s3 == 0:
                    continue
                sh = np.array([s1, s2, s3], dtype=np.float64)
                f = frac + sh
                keep = np.all((f >= -fmargin) & (f < 1 + fmargin), axis=1)
                if keep.any():
                    xs.append(w[keep] + sh @ cell)
                    owner.append(np.flatnonzero(keep))
    x = np.vstack(xs)
    owner = np.concatenate(owner)
    types = np.asarray(species, dtype=np.int32)[owner] + 1
    tree = cKDTree(x)
    rng = np.random.default_rng(seed)
    lists = tree.query_ball_point(w, list_cut)
    neigh, numneigh, offsets = [], np.zeros(n, np.int32), np.zeros(n, np.int32)
    off = 0
    for i in range(n):
        js = np.array(sorted(j for j in lists[i] if j != i), dtype=np.int32)
        if shuffle:
            js = js[rng.permutation(len(js))]
        neigh.append(js)
        numneigh[i] = len(js)
        offsets[i] = off
        off += len(js)
    maxn = max(1, int(numneigh.max()) if n else 1)
    nbr2d = np.zeros((n, maxn), np.int32)
    for i in range(n):
        nbr2d[i, :numneigh[i]] = neigh[i]
    ilist = np.arange(n, dtype=np.int32)
    if shuffle:
        ilist = ilist[rng.permutation(n)].astype(np.int32)
    return dict(x=x, type=types, owner=owner, ilist=ilist, numneigh=numneigh,
                offsets=offsets, neigh=np.concatenate(neigh) if neigh else
                np.zeros(0, np.int32), nbr2d=nbr2d, n_local=n, wrapped=w)
