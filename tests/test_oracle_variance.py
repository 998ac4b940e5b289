"""Oracle variance path: the variance map (normalized_dot_product.cpp:761-852)
and compute flare/std/atom (compute_flare_std_atom.cpp:99-213), pinned by the
reference's own cross-identity (tests/test_lammps.py:207-237): the per-atom
stds LAMMPS computes from the L^-1 + sparse-descriptor files, and from the
variance map of the power-1 model, equal the signed square roots of the GP
local variances divided by |sigma| (sparse_gp_calculator.py:94-109)."""
import numpy as np
import pytest

from flare_pp_amd.calculator import local_stds
from helpers import grouped_sparse, small_two_species

S, N, L = 2, 4, 3
SEL = [[0, 3, 5, 9, 11], [1, 2, 7, 12]]


def _setup(orc, power, sigma, seed=3, sel=SEL, cell_species_pos=None):
    cell, species, pos = cell_species_pos or small_two_species(seed=seed)
    nl = orc.NeighborList(cell, species, pos, 5.0)
    b = orc.B2Result(nl, species, S, N, L, np.full((S, S), 5.0))
    g = b.grouped()
    P, Pn, m = grouped_sparse(g, sel)
    K = orc.envs_envs(S, m, P, Pn, m, P, Pn, sigma, power)
    Li = orc.cholesky_inverse(K, 1e-8)
    var = orc.cluster_uncertainties(b, m, P, Pn, sigma, power, Li)
    blocks = np.split(g["atom_indices"], np.cumsum(g["n_by_type"])[:-1])
    want = local_stds(blocks, var, sigma)[:, 0]
    return cell, species, pos, nl, P, Pn, m, Li, var, want


def _normed(P, Pn):
    out = np.zeros_like(P)
    ok = Pn >= 1e-8
    out[ok] = P[ok] / Pn[ok, None]
    return out


@pytest.mark.parametrize("power", [1, 2])
def test_std_atom_linv_equals_gp_local_stds(orc, power):
    sigma = 1.7
    cell, species, pos, nl, P, Pn, m, Li, var, want = _setup(orc, power, sigma)
    got = orc.std_atom_linv(nl, species, S, N, L, np.full((S, S), 5.0), sigma,
                            power, m, _normed(P, Pn), Li)
    # variance scale is 1 (normalised by sigma^2); stds near a sparse point
    # are sqrt of a jitter-sized difference, so compare the variances too
    np.testing.assert_allclose(np.sign(got) * got ** 2, np.sign(want) * want ** 2,
                               atol=1e-10)
    np.testing.assert_allclose(got, want, atol=1e-5)
    assert np.abs(got).max() > 0.05


def test_std_atom_map_equals_power1_gp_local_stds(orc):
    sigma = 1.3
    cell, species, pos, nl, P, Pn, m, Li, var, want = _setup(orc, 1, sigma)
    Kinv = Li.T @ Li  # sparse_gp.cpp:588
    beta = orc.varmap_coefficients(S, m, P, Pn, sigma, Kinv)
    # sigma^2 I - sigma^4 P_hat^T K^-1 P_hat per species, symmetric
    off = np.concatenate([[0], np.cumsum(m)])
    for s in range(S):
        Ph = _normed(P, Pn)[off[s]:off[s + 1]]
        Ks = Kinv[off[s]:off[s + 1], off[s]:off[s + 1]]
        ref = sigma ** 2 * np.eye(P.shape[1]) - sigma ** 4 * Ph.T @ Ks @ Ph
        np.testing.assert_allclose(beta[s], ref, atol=1e-9 * np.abs(ref).max())
    got = orc.std_atom_map(nl, species, S, N, L, np.full((S, S), 5.0), sigma, beta)
    np.testing.assert_allclose(np.sign(got) * got ** 2, np.sign(want) * want ** 2,
                               atol=1e-8)
    assert np.abs(got).max() > 0.05


def test_varmap_self_term_needs_first_sparse_env(orc):
    # the sigma^2 diagonal is added inside the (i = 0, j = 0) iteration of the
    # non-empty loops (normalized_dot_product.cpp:829-842): a species whose
    # first sparse environment is empty gets none
    sigma = 1.1
    cell, species, pos = small_two_species(seed=4)
    # a 3x larger box with atom 0 moved 13 A away from the others: an empty
    # environment
    cell = cell * 3.0
    pos = pos.copy()
    pos[0] = [16.0, 16.0, 16.0]
    nl = orc.NeighborList(cell, species, pos, 5.0)
    b = orc.B2Result(nl, species, S, N, L, np.full((S, S), 5.0))
    g = b.grouped()
    assert (g["norms"] < 1e-8).sum() > 0
    empty = np.flatnonzero(g["norms"] < 1e-8)
    blocks = np.concatenate([[0], np.cumsum(g["n_by_type"])])
    s0 = int(np.searchsorted(blocks, empty[0], side="right") - 1)
    r0 = int(empty[0] - blocks[s0])
    full = [r for r in range(g["n_by_type"][s0]) if g["norms"][blocks[s0] + r] >= 1e-8]
    sel_a = [[r0] + full[:3] if s == s0 else [0, 1] for s in range(S)]
    sel_b = [full[:3] if s == s0 else [0, 1] for s in range(S)]
    out = []
    for sel in (sel_a, sel_b):
        P, Pn, m = grouped_sparse(g, sel)
        K = orc.envs_envs(S, m, P, Pn, m, P, Pn, sigma, 1.0)
        Li = orc.cholesky_inverse(K, 1e-8)
        out.append(orc.varmap_coefficients(S, m, P, Pn, sigma, Li.T @ Li))
    F = out[0].shape[1]
    # same non-empty environments: the maps differ by exactly sigma^2 I
    np.testing.assert_allclose(out[1][s0] - out[0][s0], sigma ** 2 * np.eye(F),
                               atol=1e-9)


def test_cluster_uncertainties_rows_subset(orc):
    """The row-subset variances (used at full size on the GPU box) are the
    full result's entries: rows of both species, one row alone."""
    cell, species, pos, nl, P, Pn, m, Li, var, want = _setup(orc, 2, 1.3)
    b = orc.B2Result(nl, species, S, N, L, np.full((S, S), 5.0))
    for rows in ([0, 5, 17, 20, 31], [len(var) - 1], list(range(len(var)))):
        sub = orc.cluster_uncertainties(b, m, P, Pn, 1.3, 2, Li, rows=rows)
        np.testing.assert_array_equal(sub, var[rows])
