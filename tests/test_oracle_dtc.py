"""Oracle self_kernel_struc (normalized_dot_product.cpp:446-578), the input of
SparseGP::predict_DTC (sparse_gp.cpp:654-680), pinned by what the mathematics
fixes: k(o, o) of an output o is the mixed second derivative of the
structure-structure energy kernel K(x, y) = sum_ij k(d_i(x), d_j(y)) along o
in both arguments (forces: positions; stress: strains, scaled by 1/V^2); its
energy entry is K itself."""
import numpy as np
import pytest

from helpers import isolated_cluster, small_two_species, strain

S, N, L = 2, 3, 2


def desc(orc, cell, species, pos, dervs=False):
    nl = orc.NeighborList(cell, species, pos, 5.0)
    b = orc.B2Result(nl, species, S, N, L, np.full((S, S), 5.0), dervs=dervs)
    return b, b.grouped()


def kk(orc, g1, g2, sigma, power):
    return orc.envs_envs(S, g1["n_by_type"], g1["descriptors"], g1["norms"],
                         g2["n_by_type"], g2["descriptors"], g2["norms"], sigma,
                         power).sum()


@pytest.mark.parametrize("power", [1, 2])
def test_self_kernel_struc_by_finite_differences(orc, power):
    cell, species, pos = small_two_species(seed=8, n_cells=1, sigma=0.15)
    sigma = 1.4
    V = orc.volume(cell)
    b, g = desc(orc, cell, species, pos, dervs=True)
    ks = orc.self_kernel_struc(b, sigma, power, V)
    n = len(species)
    # energy entry: the kernel of the structure with itself
    assert abs(ks[0] - kk(orc, g, g, sigma, power)) < 1e-12 * abs(ks[0])
    h = 2e-4

    def mixed(shift):
        vals = {}
        for sx in (1, -1):
            for sy in (1, -1):
                gx = desc(orc, *shift(sx * h))[1]
                gy = desc(orc, *shift(sy * h))[1]
                vals[sx, sy] = kk(orc, gx, gy, sigma, power)
        return (vals[1, 1] - vals[1, -1] - vals[-1, 1] + vals[-1, -1]) / (4 * h * h)

    for a, m in [(0, 0), (1, 1), (n - 1, 2), (2, 0)]:
        def move(d, a=a, m=m):
            p = pos.copy()
            p[a, m] += d
            return cell, species, p
        want = mixed(move)
        got = ks[1 + 3 * a + m]
        assert abs(got - want) < 2e-5 * max(abs(want), 1e-3), (a, m, got, want)
    comps = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for q, (i, j) in enumerate(comps):
        def deform(d, i=i, j=j):
            c2, p2 = strain(cell, pos, j, i, d)
            return c2, species, p2
        want = mixed(deform) / V ** 2
        got = ks[1 + 3 * n + q]
        assert abs(got - want) < 2e-5 * max(abs(want), 1e-9), (q, got, want)


def test_self_kernel_struc_empty_atoms(orc):
    # isolated atoms (no neighbors) contribute nothing
    cell, species, pos = isolated_cluster(seed=1, n=5)
    pos = np.vstack([pos, [2.0, 2.0, 2.0]])  # far from the cluster
    species = np.append(species, 0).astype(np.int32)
    b, g = desc(orc, cell, species, pos, dervs=True)
    ks = orc.self_kernel_struc(b, 1.0, 2.0, orc.volume(cell))
    n = len(species)
    assert np.all(ks[1 + 3 * (n - 1):1 + 3 * n] == 0)
    assert abs(ks[0] - kk(orc, g, g, 1.0, 2.0)) < 1e-12 * abs(ks[0])
