"""Oracle B2, kernels, sparse-GP prediction and mapped pair style, pinned by
invariances, finite differences and the reference's cross-identities
(b2.cpp; normalized_dot_product.cpp; sparse_gp.cpp; pair_flare.cpp;
ctests/test_kernels.cpp; tests/test_lammps.py)."""
import numpy as np
import pytest

from conftest import random_rotation
from helpers import grouped_sparse, isolated_cluster, small_two_species, strain

S2 = dict(S=2, N=4, L=3)


def b2_of(orc, cell, species, pos, S, N, L, rc=5.0, dervs=False,
          cutoff="quadratic", cutoffs=None, method="cells"):
    nl = orc.NeighborList(cell, species, pos, rc if cutoffs is None else
                          np.max(cutoffs), method)
    cut = np.full((S, S), rc) if cutoffs is None else cutoffs
    return orc.B2Result(nl, species, S, N, L, cut, cutoff=cutoff, dervs=dervs)


def test_b2_sizes(orc, cfg1):
    c = cfg1
    b = b2_of(orc, c["cell"], c["species"], c["positions"], 1, 8, 3)
    assert b.F == 144 and b.W == 128  # F = nr(nr+1)/2 (L+1)


def test_b2_contraction_explicit(orc):
    cell, species, pos = small_two_species()
    b = b2_of(orc, cell, species, pos, **S2)
    A, B, nrm = b.atom_order()
    N, L, S = 4, 3, 2
    nr, H = S * N, (L + 1) ** 2
    A3 = A.reshape(-1, nr, H)
    want = []
    for n1 in range(nr):
        for n2 in range(n1, nr):
            for l in range(L + 1):
                sl = slice(l * l, (l + 1) ** 2)
                want.append(np.einsum("im,im->i", A3[:, n1, sl], A3[:, n2, sl]))
    np.testing.assert_allclose(B, np.stack(want, 1), rtol=1e-13, atol=1e-13)
    np.testing.assert_allclose(nrm, np.linalg.norm(B, axis=1), rtol=1e-14)


def test_single_bond_layout_independent_build(orc):
    """The single-bond array A (b2.cpp:244-399) built independently in numpy
    from closed forms: all periodic images by brute force, the pair kept when
    0 < r <= rc(s_i, s_j) (row = centre species, asymmetric here), g_n =
    T_n(r / rc) (r - rc)^2 with T_n(cos t) = cos(n t), real Y_lm of l <= 3
    written out, accumulated into channel s_j N + n and lm = l^2 + m + l.
    A transposed cutoff matrix, a channel taken from the centre's species or
    a shifted lm index each fail it."""
    from test_oracle_math import closed_forms
    S, N, L = 2, 5, 3
    cut = np.array([[5.0, 4.4], [4.0, 4.7]])
    cell, species, pos = small_two_species(seed=21)
    b = b2_of(orc, cell, species, pos, S, N, L, cutoffs=cut)
    A_o = b.atom_order()[0]
    H = (L + 1) ** 2
    A = np.zeros((len(species), S * N * H))
    shifts = np.array([[a, c, d] for a in range(-2, 3) for c in range(-2, 3)
                       for d in range(-2, 3)]) @ cell
    for i in range(len(species)):
        for j in range(len(species)):
            rc = cut[species[i], species[j]]
            for dv in pos[j] - pos[i] + shifts:
                r = np.linalg.norm(dv)
                if not 0 < r <= rc:
                    continue
                g = np.cos(np.arange(N) * np.arccos(r / rc)) * (r - rc) ** 2
                y = closed_forms(*dv)
                for n in range(N):
                    for (l, m), v in y.items():
                        A[i, (species[j] * N + n) * H + l * l + m + l] += g[n] * v
    assert np.abs(A - A_o).max() < 1e-13 * np.abs(A).max()


def test_b2_rotation_translation_permutation(orc):
    rng = np.random.default_rng(4)
    cell, species, pos = small_two_species(seed=1)
    B0 = b2_of(orc, cell, species, pos, **S2).atom_order()[1]
    R = random_rotation(rng)
    B1 = b2_of(orc, cell @ R.T, species, pos @ R.T, **S2).atom_order()[1]
    np.testing.assert_allclose(B1, B0, rtol=1e-10, atol=1e-10 * np.abs(B0).max())
    B2 = b2_of(orc, cell, species, pos + np.array([0.3, -1.7, 2.9]),
               **S2).atom_order()[1]
    np.testing.assert_allclose(B2, B0, rtol=1e-10, atol=1e-10 * np.abs(B0).max())
    perm = rng.permutation(len(species))
    B3 = b2_of(orc, cell, species[perm], pos[perm], **S2).atom_order()[1]
    np.testing.assert_allclose(B3, B0[perm], rtol=1e-10, atol=1e-10 * np.abs(B0).max())


@pytest.mark.parametrize("cutoff", ["quadratic", "cosine"])
def test_b2_force_dervs_fd(orc, cutoff):
    # dB_i/dr_ij against central differences in x_j (b2.cpp:219-228)
    cell, species, pos = isolated_cluster(seed=3)
    b = b2_of(orc, cell, species, pos, **S2, dervs=True, cutoff=cutoff)
    g = b.grouped()
    h = 1e-5
    n = len(species)
    base_rows = {a: r for r, a in enumerate(g["atom_indices"])}
    pair = 0
    for r in range(n):
        i = g["atom_indices"][r]
        for k in range(g["neighbor_counts"][r]):
            j = g["neighbor_indices"][pair]
            for c in range(3):
                pp, pm = pos.copy(), pos.copy()
                pp[j, c] += h
                pm[j, c] -= h
                Bp = b2_of(orc, cell, species, pp, **S2, cutoff=cutoff).atom_order()[1]
                Bm = b2_of(orc, cell, species, pm, **S2, cutoff=cutoff).atom_order()[1]
                fd = (Bp[i] - Bm[i]) / (2 * h)
                an = g["force_dervs"][3 * pair + c]
                np.testing.assert_allclose(an, fd, atol=1e-6 * max(1, np.abs(fd).max()))
                np.testing.assert_allclose(g["force_dots"][3 * pair + c],
                                           an @ g["descriptors"][base_rows[i]],
                                           rtol=1e-10, atol=1e-12)
            pair += 1
    assert pair == b.n_pairs > 0


def _gp_setup(orc, cell, species, pos, S, N, L, sel, dervs=True):
    b = b2_of(orc, cell, species, pos, S, N, L, dervs=dervs)
    g = b.grouped()
    P, Pn, m = grouped_sparse(g, sel)
    return b, g, P, Pn, m


SEL = [[0, 3, 5, 9], [1, 2, 7]]


def test_envs_struc_relations(orc):
    cell, species, pos = small_two_species(seed=2)
    sigma, power = 1.7, 2.0
    b, g, P, Pn, m = _gp_setup(orc, cell, species, pos, **S2, sel=SEL)
    V = orc.volume(cell)
    K = orc.envs_struc(b, m, P, Pn, sigma, power, V)
    # energy column = sum over atoms of envs_envs (test_kernels.cpp:55-75)
    ee = orc.envs_envs(2, m, P, Pn, g["n_by_type"], g["descriptors"], g["norms"],
                       sigma, power)
    np.testing.assert_allclose(K[:, 0], ee.sum(1), rtol=1e-12)
    # force columns = -d(energy column)/dx (test_kernels.cpp:140-406)
    h = 1e-5
    n = len(species)
    for a, c in [(0, 0), (5, 2), (n - 1, 1)]:
        pp, pm = pos.copy(), pos.copy()
        pp[a, c] += h
        pm[a, c] -= h
        Ep = orc.envs_struc(b2_of(orc, cell, species, pp, **S2, dervs=True), m, P,
                            Pn, sigma, power, V)[:, 0]
        Em = orc.envs_struc(b2_of(orc, cell, species, pm, **S2, dervs=True), m, P,
                            Pn, sigma, power, V)[:, 0]
        np.testing.assert_allclose(K[:, 1 + 3 * a + c], -(Ep - Em) / (2 * h),
                                   atol=1e-7 * np.abs(K[:, 0]).max())
    # stress columns = -(1/V) dE/d(strain) (xx, xy, xz, yy, yz, zz)
    comps = [(0, 0), (0, 1), (0, 2), (1, 1), (1, 2), (2, 2)]
    for q, (i, j) in enumerate(comps):
        cp, xp = strain(cell, pos, j, i, h)
        cm, xm = strain(cell, pos, j, i, -h)
        Ep = orc.envs_struc(b2_of(orc, cp, species, xp, **S2, dervs=True), m, P,
                            Pn, sigma, power, V)[:, 0]
        Em = orc.envs_struc(b2_of(orc, cm, species, xm, **S2, dervs=True), m, P,
                            Pn, sigma, power, V)[:, 0]
        np.testing.assert_allclose(K[:, 1 + 3 * n + q], -(Ep - Em) / (2 * h) / V,
                                   atol=1e-7 * np.abs(K[:, 0]).max() / V)
    # forces sum to zero (translation invariance)
    np.testing.assert_allclose(K[:, 1:1 + 3 * n].reshape(len(P), n, 3).sum(1), 0,
                               atol=1e-10 * np.abs(K).max())


def test_envs_envs_diagonal_is_sigma2(orc, cfg1):
    c = cfg1
    b = b2_of(orc, c["cell"], c["species"], c["positions"], 1, 8, 3)
    g = b.grouped()
    k = orc.envs_envs(1, g["n_by_type"], g["descriptors"], g["norms"],
                      g["n_by_type"], g["descriptors"], g["norms"], 2.0, 2.0)
    np.testing.assert_allclose(np.diag(k), 4.0, rtol=1e-14)
    np.testing.assert_allclose(k, k.T, rtol=1e-13)
    assert k.max() <= 4.0 * (1 + 1e-14)


def test_cholesky_inverse_and_variance(orc, cfg1):
    c = cfg1
    b = b2_of(orc, c["cell"], c["species"], c["positions"], 1, 8, 3)
    g = b.grouped()
    P, Pn, m = grouped_sparse(g, [c["sparse_atoms"][:20]])
    K = orc.envs_envs(1, m, P, Pn, m, P, Pn, 2.0, 2.0)
    jit = 1e-3
    Li = orc.cholesky_inverse(K, jit)
    Lm = np.linalg.inv(Li)
    np.testing.assert_allclose(Lm @ Lm.T, K + jit * np.eye(len(K)), atol=1e-10)
    assert np.allclose(np.triu(Li, 1), 0)
    var = orc.cluster_uncertainties(b, m, P, Pn, 2.0, 2.0, Li)
    # independent numpy evaluation of sigma^2 - k^T (K + jI)^-1 k
    kv = orc.envs_envs(1, g["n_by_type"], g["descriptors"], g["norms"], m, P, Pn,
                       2.0, 2.0)
    want = 4.0 - np.einsum("iu,iu->i", kv @ np.linalg.inv(K + jit * np.eye(len(K))), kv)
    np.testing.assert_allclose(var, want, atol=1e-9)
    assert var.min() > -1e-9  # test_sparse_gp.cpp:86-100
    # at a sparse environment u the variance is j e_u^T K (K + jI)^-1 e_u,
    # which lies in [0, j]
    sel = np.asarray(c["sparse_atoms"][:20])
    rows = np.searchsorted(g["atom_indices"], sel)
    assert var[rows].min() > -1e-12 and var[rows].max() <= jit * (1 + 1e-9)


@pytest.mark.parametrize("power", [1, 2])
def test_mapped_pair_equals_gp_mean(orc, power):
    # tests/test_lammps.py:97-102: pair_flare with the mapping coefficients
    # reproduces the GP mean energy, forces and stress
    cell, species, pos = small_two_species(seed=5)
    sigma = 1.3
    S, N, L = 2, 4, 3
    b, g, P, Pn, m = _gp_setup(orc, cell, species, pos, S, N, L, SEL)
    alpha = np.random.default_rng(9).normal(size=len(P))
    V = orc.volume(cell)
    efs = orc.predict_mean(b, m, P, Pn, sigma, power, V, alpha)
    beta = orc.mapping_coefficients(S, m, P, Pn, sigma, power, alpha)
    nl = orc.NeighborList(cell, species, pos, 5.0)
    E, F, vir = orc.pair_compute(nl, species, S, N, L, power, np.full((2, 2), 5.0),
                                 beta)
    n = len(species)
    scale = np.abs(F).max()
    assert abs(E - efs[0]) < 1e-12 * max(1, abs(E))
    np.testing.assert_allclose(F, efs[1:1 + 3 * n].reshape(n, 3), atol=1e-11 * scale)
    s = efs[1 + 3 * n:]  # xx xy xz yy yz zz
    lmp = np.array([s[0], s[3], s[5], s[1], s[2], s[4]]) * V
    np.testing.assert_allclose(vir, lmp, atol=1e-11 * np.abs(lmp).max())


@pytest.mark.parametrize("cutoff", ["quadratic", "cosine"])
def test_pair_forces_and_virial_fd(orc, cutoff):
    cell, species, pos = small_two_species(seed=6)
    S, N, L = 2, 4, 3
    F_ = (S * N) * (S * N + 1) // 2 * (L + 1)
    rng = np.random.default_rng(1)
    A = rng.normal(0, 1e-2, (S, F_, F_))
    beta = 0.5 * (A + np.transpose(A, (0, 2, 1)))
    cut = np.array([[5.0, 4.6], [4.6, 4.2]])

    def ev(c, x):
        nl = orc.NeighborList(c, species, x, 5.0)
        return orc.pair_compute(nl, species, S, N, L, 2, cut, beta, cutoff=cutoff,
                                packed=False)

    E, F, vir = ev(cell, pos)
    h = 1e-5
    for a, c in [(0, 0), (3, 1), (9, 2)]:
        pp, pm = pos.copy(), pos.copy()
        pp[a, c] += h
        pm[a, c] -= h
        fd = -(ev(cell, pp)[0] - ev(cell, pm)[0]) / (2 * h)
        assert abs(fd - F[a, c]) < 1e-7 * max(1, np.abs(F).max())
    np.testing.assert_allclose(F.sum(0), 0, atol=1e-10 * np.abs(F).max())
    # virial W_ab = -dE/d(strain_ab) for the diagonal (LAMMPS xx yy zz)
    for q in range(3):
        cp, xp = strain(cell, pos, q, q, h)
        cm, xm = strain(cell, pos, q, q, -h)
        fd = -(ev(cp, xp)[0] - ev(cm, xm)[0]) / (2 * h)
        assert abs(fd - vir[q]) < 1e-6 * max(1, np.abs(vir).max())


def test_pair_empty_environment(orc):
    # an isolated atom has no energy and no force (pair_flare.cpp:135-136)
    cell = np.eye(3) * 30.0
    pos = np.array([[1.0, 1.0, 1.0], [15.0, 15.0, 15.0], [17.0, 15.5, 15.0]])
    species = np.array([0, 1, 0], dtype=np.int32)
    S, N, L = 2, 3, 2
    F_ = (S * N) * (S * N + 1) // 2 * (L + 1)
    beta = np.tile(np.eye(F_) * 0.01, (S, 1, 1))
    nl = orc.NeighborList(cell, species, pos, 5.0)
    E, F, vir, ea = orc.pair_compute(nl, species, S, N, L, 2, np.full((2, 2), 5.0),
                                     beta, packed=False, want_eatom=True)
    assert ea[0] == 0 and np.all(F[0] == 0)
    assert np.isfinite(E) and ea[1] != 0


@pytest.mark.parametrize("power", [1, 2])
def test_pair_on_lammps_ghost_list_equals_periodic(orc, power):
    # PairFLARE::compute on a full list with ghost atoms, ghost forces summed
    # onto their owners (LAMMPS' reverse communication, newton on), equals the
    # periodic evaluation on the Structure list
    from helpers import lammps_ghost_list
    cell, species, pos = small_two_species(seed=6)
    S, N, L = 2, 4, 3
    cut = np.array([[5.0, 4.6], [4.6, 4.2]])
    nr = S * N
    F = nr * (nr + 1) // 2 * (L + 1)
    rng = np.random.default_rng(2)
    if power == 2:
        a = rng.normal(0, 1e-3, (S, F, F))
        beta = (a + a.transpose(0, 2, 1)) / 2
    else:
        beta = rng.normal(0, 1e-2, (S, F))
    g = lammps_ghost_list(cell, species, pos, 6.0, 6.0, seed=3)
    e, f_all, vir, ea = orc.pair_compute_lmp(g["x"], g["type"], g["ilist"],
                                             g["numneigh"], g["offsets"], g["neigh"],
                                             S, N, L, power, cut, beta)
    n = len(species)
    f = np.zeros((n, 3))
    np.add.at(f, g["owner"], f_all)
    nl = orc.NeighborList(cell, species, pos, cut.max())
    E0, F0, V0, ea0 = orc.pair_compute(nl, species, S, N, L, power, cut, beta,
                                       packed=False, want_eatom=True)
    scale = np.abs(F0).max()
    assert abs(e - E0) < 1e-12 * max(1.0, abs(E0))
    np.testing.assert_allclose(f, F0, atol=1e-11 * scale)
    np.testing.assert_allclose(vir, V0, atol=1e-11 * np.abs(V0).max())
    np.testing.assert_allclose(ea[:n], ea0, atol=1e-12 * np.abs(ea0).max())
    assert np.all(ea[n:] == 0)
