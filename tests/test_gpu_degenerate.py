"""Degenerate inputs on the GPU path: structures without atoms, a LAMMPS
rank without local atoms (pair_flare.cpp:68-184 loops over inum centres, so
an empty rank contributes nothing), and structures that lack some of the
potential's species (empty per-species blocks of the DescriptorValues
layout, b2.cpp:76-179; sparse_gp.cpp:612-628 with zero rows of a type)."""
import numpy as np
import pytest
import torch

from flare_pp_amd import (B2, NormalizedDotProduct, PairFLARE, SparseGP,
                          Structure, pack_beta, synth)
from helpers import small_two_species

pytestmark = pytest.mark.gpu
DEV = "cuda:0"
FORMS = ("plain", "adjoint_fp64", "adjoint_fp32", "adjoint_mixed")


def make_pair(S, N, L, seed):
    F = (S * N) * (S * N + 1) // 2 * (L + 1)
    beta = synth.symmetric_beta(S, F, 1e-3, seed)
    return PairFLARE.from_coefficients(2, S, N, L, np.full((S, S), 5.0),
                                       np.stack([pack_beta(b) for b in beta]))


def nrel(a, b):
    return np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300)


def test_structure_without_atoms():
    print("structure, no atoms", flush=True)
    b2 = B2("chebyshev", "quadratic", [0, 5.0], [], [2, 8, 3])
    s = Structure(np.eye(3) * 10.0, np.zeros(0, np.int32), np.zeros((0, 3)), 5.0, [b2])
    assert s.n_neighbors == 0
    assert len(s.neighbor_count) == 0
    assert list(s.cumulative_neighbor_count) == [0]
    d = s.descriptors[0]
    assert d.n_atoms == 0 and list(d.n_clusters_by_type) == [0, 0]


@pytest.mark.parametrize("form", FORMS)
def test_pair_without_atoms(form):
    # energy and virial written as zeros; a following call on a real
    # structure is unaffected
    print(f"pair, no atoms, form {form}", flush=True)
    pair = make_pair(2, 8, 3, seed=1)
    cell, species, pos = small_two_species(seed=3)
    e0, f0, v0 = pair.compute(cell, species, pos, form=form)
    e, f, v, ea = pair.compute(np.eye(3) * 12.0, np.zeros(0, np.int32),
                               np.zeros((0, 3)), form=form, want_eatom=True)
    assert e == 0.0 and f.shape == (0, 3) and ea.shape == (0,)
    assert not np.any(v)
    e1, f1, v1 = pair.compute(cell, species, pos, form=form)
    assert e1 == e0 and np.array_equal(f1, f0) and np.array_equal(v1, v0)


@pytest.mark.parametrize("case", ["first_of_two", "middle_of_three"])
def test_pair_with_absent_species(orc, case):
    # a 2-species potential on a species-1-only crystal (species 0's rows and
    # GEMM block empty), and a 3-species potential (24 channels) on species
    # 0 and 2 only (an empty block between two full ones)
    if case == "first_of_two":
        pair = make_pair(2, 8, 3, seed=5)
        cell, species, pos = synth.fcc(2, 4.05, 0.1, [0, 32], 7)
    else:
        pair = make_pair(3, 8, 3, seed=6)
        cell, species, pos = synth.fcc(2, 4.05, 0.1, [16, 0, 16], 8)
    nl = orc.NeighborList(cell, species, pos, pair.cutoff)
    e0, f0, v0, ea0 = orc.pair_compute(nl, species, pair.n_species, pair.n_max,
                                       pair.l_max, pair.power, pair.cutoff_matrix,
                                       pair.beta, cutoff=pair.cutoff_function,
                                       packed=False, want_eatom=True)
    for form in FORMS:
        print(f"absent species ({case}), form {form}", flush=True)
        e, f, v, ea = pair.compute(cell, species, pos, form=form, want_eatom=True)
        tol = 1e-10 if form in ("plain", "adjoint_fp64") else 1e-3
        assert abs(e - e0) <= tol * np.abs(ea0).sum(), form
        assert nrel(f, f0) < tol, form
        assert nrel(v, v0) < tol, form
        assert nrel(ea, ea0) < tol, form


@pytest.mark.parametrize("form", ["plain", "adjoint_fp64", "adjoint_mixed"])
def test_neigh_list_without_local_atoms(form):
    # inum = 0 with ghost atoms present: energy and virial written as zeros,
    # the accumulated forces and per-atom energies left as they were
    print(f"LAMMPS list, inum 0, form {form}", flush=True)
    pair = make_pair(2, 8, 3, seed=2)
    n_all = 5
    rng = np.random.default_rng(4)
    x = torch.as_tensor(rng.uniform(0, 6, (n_all, 3)), dtype=torch.float64, device=DEV)
    typ = torch.ones(n_all, dtype=torch.int32, device=DEV)
    empty = torch.zeros(0, dtype=torch.int32, device=DEV)
    e = torch.full((1,), 7.0, dtype=torch.float64, device=DEV)
    v = torch.full((6,), 7.0, dtype=torch.float64, device=DEV)
    f_in = torch.as_tensor(rng.normal(size=(n_all, 3)), dtype=torch.float64, device=DEV)
    ea_in = torch.as_tensor(rng.normal(size=n_all), dtype=torch.float64, device=DEV)
    f, ea = f_in.clone(), ea_in.clone()
    pair.compute_neigh_device(n_all, 0, empty, empty, empty, 1, 1, x, typ, form,
                              e, f, v, ea)
    torch.cuda.synchronize()
    assert e.item() == 0.0 and not torch.any(v)
    assert torch.equal(f, f_in) and torch.equal(ea, ea_in)


def test_gp_prediction_without_atoms():
    # a model trained on a real structure, asked about an empty one: the
    # energy, the six stress components and the (empty) variances
    print("sparse GP, no atoms", flush=True)
    cell, species, pos = small_two_species(seed=9)
    b2 = B2("chebyshev", "quadratic", [0, 5.0], [], [2, 8, 3])
    s = Structure(cell, species, pos, 5.0, [b2])
    gp = SparseGP([NormalizedDotProduct(2.0, 2)], 0.1, 0.1, 0.1)
    gp.add_specific_environments(s, [0, 1, 2, 17, 20])
    gp.set_alpha(np.random.default_rng(1).normal(size=5))
    empty = Structure(np.eye(3) * 11.0, np.zeros(0, np.int32), np.zeros((0, 3)), 5.0,
                      [b2])
    for form in FORMS:
        efs = gp.predict_mean(empty, form=form)
        assert efs.shape == (7,) and not np.any(efs), form
    efs, (var,) = gp.predict_local_uncertainties(empty)
    assert efs.shape == (7,) and not np.any(efs) and var.shape == (0,)
    # and the model still predicts the real structure afterwards
    a = gp.predict_mean(s, form="adjoint_fp64")
    b = gp.predict_mean(s, form="adjoint_fp64")
    assert np.array_equal(a, b)
