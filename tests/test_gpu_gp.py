"""GPU sparse-GP prediction against the oracle (sparse_gp.cpp:93-133,
576-628, 682-702; normalized_dot_product.cpp:27-210, 643-759)."""
import numpy as np
import pytest

from flare_pp_amd import (B2, NormalizedDotProduct, PairFLARE, SparseGP,
                          Structure, efs_to_results, local_stds, synth)
from helpers import grouped_sparse, small_two_species

pytestmark = pytest.mark.gpu


def nrel(a, b):
    return np.abs(np.asarray(a) - b).max() / max(np.abs(b).max(), 1e-300)


def variance_bound(Kuu, kvec, sigma, jitter=1e-8):
    """Per-row rounding bound of v = k_self - |L^-1 k|^2 (sparse_gp.cpp:
    93-133) for two FP64 evaluations: a backward-stable Cholesky gives
    L L^T = K + jitter I + dK with |dK| <= M eps |K|, which moves k^T (K +
    jitter I)^-1 k by |w|^2 |dK|, w = (K + jitter I)^-1 k; the rest rounds at
    eps sigma^2.  The conditioning of K_uu enters through |w|."""
    from scipy.linalg import cho_factor, cho_solve
    M = len(Kuu)
    eps = np.finfo(np.float64).eps
    Kj = Kuu + jitter * np.eye(M)
    w = cho_solve(cho_factor(Kj, lower=True), np.asarray(kvec).T)
    return 2 * (M * eps * np.linalg.norm(Kj, 2) * np.sum(w ** 2, axis=0)
                + 16 * eps * sigma ** 2)


def build(cell, species, pos, settings, rc, sigma, power, atoms, alpha):
    b2 = B2("chebyshev", "quadratic", [0, rc], [], settings)
    s = Structure(cell, species, pos, rc, [b2])
    gp = SparseGP([NormalizedDotProduct(sigma, power)], 0.1, 0.1, 0.1)
    gp.add_specific_environments(s, atoms)
    gp.set_alpha(alpha)
    return s, gp, b2


def oracle_side(orc, cell, species, pos, settings, rc, atoms):
    nl = orc.NeighborList(cell, species, pos, rc)
    S = settings[0]
    o = orc.B2Result(nl, species, *settings, np.full((S, S), rc), dervs=True)
    g = o.grouped()
    sel = [[k for k, a in enumerate(g["atom_indices"][off:off + cnt]) if a in atoms]
           for off, cnt in zip(np.concatenate([[0], np.cumsum(g["n_by_type"])[:-1]]),
                               g["n_by_type"])]
    P, Pn, m = grouped_sparse(g, sel)
    return o, g, P, Pn, m


@pytest.mark.parametrize("power", [1, 2])
def test_config1_mean_and_variance(orc, cfg1, power):
    c = cfg1
    s, gp, b2 = build(c["cell"], c["species"], c["positions"], c["settings"], 5.0,
                      c["sigma"], power, c["sparse_atoms"], c["alpha"])
    o, g, P, Pn, m = oracle_side(orc, c["cell"], c["species"], c["positions"],
                                 c["settings"], 5.0, c["sparse_atoms"])
    V = orc.volume(c["cell"])
    want = orc.predict_mean(o, m, P, Pn, c["sigma"], power, V, c["alpha"])
    n = s.noa
    for form in ("plain", "adjoint_fp64"):
        print(f"form {form} power {power}", flush=True)
        efs = gp.predict_mean(s, form=form)
        assert abs(efs[0] - want[0]) < 1e-10 * max(1, abs(want[0]))
        assert nrel(efs[1:1 + 3 * n], want[1:1 + 3 * n]) < 1e-10
        assert nrel(efs[1 + 3 * n:], want[1 + 3 * n:]) < 1e-10
    # variance: compared relative to sigma^2, the scale the cancellation
    # sigma^2 - |L^-1 k|^2 is measured on
    Li_o = orc.cholesky_inverse(orc.envs_envs(1, m, P, Pn, m, P, Pn, c["sigma"],
                                              power), 1e-8)
    var_o = orc.cluster_uncertainties(o, m, P, Pn, c["sigma"], power, Li_o)
    _, (var,) = gp.predict_local_uncertainties(s)
    k_rows = orc.envs_envs(1, g["n_by_type"], g["descriptors"], g["norms"], m, P, Pn,
                           c["sigma"], power)
    bound = variance_bound(gp.Kuu, k_rows, c["sigma"])
    err = np.abs(var - var_o)
    print(f"power {power}: variance error max {err.max():.2e} sigma^2 = "
          f"{c['sigma'] ** 2}, bound max {bound.max():.2e} median "
          f"{np.median(bound):.2e}; cond(K_uu + jitter) "
          f"{np.linalg.cond(gp.Kuu + 1e-8 * np.eye(len(gp.Kuu))):.2e}", flush=True)
    assert np.all(err <= bound)
    Kuu = gp.Kuu
    K_o = orc.envs_envs(1, m, P, Pn, m, P, Pn, c["sigma"], power)
    assert nrel(Kuu, K_o) < 1e-12
    Li = gp.L_inv
    Lm = np.linalg.inv(Li)
    assert nrel(Lm @ Lm.T, K_o + 1e-8 * np.eye(len(K_o))) < 1e-10
    # ASE-side conventions (sparse_gp_calculator.py:56-109)
    e, f, st = efs_to_results(efs, n)
    assert e == efs[0] and f.shape == (n, 3) and st[0] == -efs[1 + 3 * n]
    stds = local_stds(s.descriptors[0].atom_indices, var, c["sigma"])
    assert stds.shape == (n, 3)


def test_two_species_mean_variance_mapping(orc, tmp_path):
    cell, species, pos = small_two_species(seed=12)
    settings, rc, sigma = [2, 4, 3], 5.0, 1.5
    atoms = [0, 2, 3, 5, 8, 11, 13, 20, 27]
    alpha = np.random.default_rng(3).normal(size=len(atoms))
    s, gp, b2 = build(cell, species, pos, settings, rc, sigma, 2, atoms, alpha)
    o, g, P, Pn, m = oracle_side(orc, cell, species, pos, settings, rc, atoms)
    V = orc.volume(cell)
    alpha_grouped = np.concatenate([alpha[[atoms.index(a) for a in blk if a in atoms]]
                                    for blk in s.descriptors[0].atom_indices])
    gp.set_alpha(alpha_grouped)
    want = orc.predict_mean(o, m, P, Pn, sigma, 2, V, alpha_grouped)
    efs = gp.predict_mean(s, form="adjoint_fp64")
    assert nrel(efs, want) < 1e-10
    beta = gp.mapping_coefficients()
    beta_o = orc.mapping_coefficients(2, m, P, Pn, sigma, 2, alpha_grouped)
    assert nrel(beta, beta_o) < 1e-12
    # GP -> coefficient file -> mapped pair style reproduces the GP mean
    path = tmp_path / "lmp.flare"
    gp.write_mapping_coefficients(str(path), "test", 0)
    pair = PairFLARE(str(path))
    e, f, v = pair.compute(cell, species, pos, form="plain")
    n = len(species)
    assert abs(e - efs[0]) < 1e-10 * max(1, abs(e))
    assert nrel(f.ravel(), efs[1:1 + 3 * n]) < 1e-10
    st = efs[1 + 3 * n:]
    assert nrel(v, np.array([st[0], st[3], st[5], st[1], st[2], st[4]]) * V) < 1e-9


def test_envs_envs(orc, cfg1):
    c = cfg1
    s, gp, b2 = build(c["cell"], c["species"], c["positions"], c["settings"], 5.0,
                      2.0, 2, c["sparse_atoms"], c["alpha"])
    k = NormalizedDotProduct(2.0, 2)
    from flare_pp_amd import ClusterDescriptor
    envs = ClusterDescriptor(s.descriptors[0])
    K = k.envs_envs(envs, gp.sparse_descriptors[0])
    o, g, P, Pn, m = oracle_side(orc, c["cell"], c["species"], c["positions"],
                                 c["settings"], 5.0, c["sparse_atoms"])
    want = orc.envs_envs(1, g["n_by_type"], g["descriptors"], g["norms"], m, P, Pn,
                         2.0, 2)
    assert nrel(K, want) < 1e-12


def test_config3_mean(orc):
    """Config 3 (32,000 atoms, 500 sparse) at full size against the oracle:
    for power 2 the GP mean equals the mapped pair energy of
    beta_s = sigma^2 sum_u alpha_u p_u p_u^T (tests/test_lammps.py:97-102),
    so the oracle's pair_flare restatement on the GPU-made coefficients
    (themselves pinned to the oracle's mapping at config 1 size) gives E, F
    and the stress of the same structure."""
    c = synth.config3()
    s, gp, b2 = build(c["cell"], c["species"], c["positions"], c["settings"], 5.0,
                      c["sigma"], c["power"], c["sparse_atoms"], c["alpha"])
    e_plain = gp.predict_mean(s, form="plain").copy()
    e_adj = gp.predict_mean(s, form="adjoint_fp64")
    assert nrel(e_adj, e_plain) < 1e-11
    packed = gp.mapping_coefficients()
    nl = orc.NeighborList(c["cell"], c["species"], c["positions"], 5.0)
    e0, f0, v0, ea0 = orc.pair_compute(nl, c["species"], 2, 8, 4, 2,
                                       np.full((2, 2), 5.0), packed, want_eatom=True)
    n = s.noa
    st = e_plain[1 + 3 * n:]
    V = orc.volume(c["cell"])
    errs = (abs(e_plain[0] - e0) / np.abs(ea0).sum(), nrel(e_plain[1:1 + 3 * n], f0.ravel()),
            nrel(np.array([st[0], st[3], st[5], st[1], st[2], st[4]]) * V, v0))
    print("config 3 GP mean vs oracle mapped: energy rel %.2e force rel %.2e "
          "stress rel %.2e" % errs, flush=True)
    assert errs[0] < 1e-10 and errs[1] < 1e-10 and errs[2] < 1e-10


def test_config4_variances_oracle_subset(orc):
    """Config 4 at full size (55,296 atoms, 3 species, 2,000 sparse): the
    FP64 local variances of 256 seeded atoms against the oracle's
    compute_cluster_uncertainties (sparse_gp.cpp:93-133) on the same rows,
    within the conditioning bound of each row."""
    c = synth.config4()
    s, gp, b2 = build(c["cell"], c["species"], c["positions"], c["settings"],
                      c["cutoff"], c["sigma"], c["power"], c["sparse_atoms"],
                      c["alpha"])
    _, (var,) = gp.predict_local_uncertainties(s, form="plain")
    var = var.copy()
    S = c["settings"][0]
    nl = orc.NeighborList(c["cell"], c["species"], c["positions"], c["cutoff"])
    o = orc.B2Result(nl, c["species"], *c["settings"], np.full((S, S), c["cutoff"]))
    g = o.grouped()
    off = np.concatenate([[0], np.cumsum(g["n_by_type"])])
    sel = [[k for k, a in enumerate(g["atom_indices"][off[t]:off[t + 1]])
            if a in set(c["sparse_atoms"])] for t in range(S)]
    P, Pn, m = grouped_sparse(g, sel)
    K_o = orc.envs_envs(S, m, P, Pn, m, P, Pn, c["sigma"], c["power"])
    assert nrel(gp.Kuu, K_o) < 1e-12
    Li_o = orc.cholesky_inverse(K_o, 1e-8)
    rows = np.sort(np.random.default_rng(41).choice(s.noa, 256, replace=False))
    var_o = orc.cluster_uncertainties(o, m, P, Pn, c["sigma"], c["power"], Li_o,
                                      rows=rows)
    ns = [int(np.sum((rows >= off[t]) & (rows < off[t + 1]))) for t in range(S)]
    k_rows = orc.envs_envs(S, ns, g["descriptors"][rows], g["norms"][rows], m, P, Pn,
                           c["sigma"], c["power"])
    bound = variance_bound(K_o, k_rows, c["sigma"])
    err = np.abs(var[rows] - var_o)
    print(f"config 4: variance error max {err.max():.2e} (sigma^2 = "
          f"{c['sigma'] ** 2}), bound min {bound.min():.2e} median "
          f"{np.median(bound):.2e}; variance/sigma^2 median "
          f"{np.median(var_o) / c['sigma'] ** 2:.2e}", flush=True)
    assert np.all(err <= bound)


# Low-precision forms.  FP32: t = d . p_hat in FP32; mixed: t and G = c P_hat
# on FP16 MFMA with hi + lo splitting, power-of-two row scaling and FP32
# accumulation.  k, c, e, w, the Cholesky factor and the variance step stay
# FP64.  Tolerances below are the stated error budget of each form;
# observed errors are printed.
LOWP_TOL = {
    # form: (energy rel, force rel to max |F|, variance abs / sigma^2).  The
    # variance k_ii - |L^-1 k_i|^2 cancels: an error dt in t = d . p_hat
    # reaches it multiplied by |L^-1|, so its budget is the loosest.
    "adjoint_fp32": (1e-5, 1e-4, 5e-3),
    "adjoint_mixed": (1e-5, 1e-4, 5e-3),
}


@pytest.mark.parametrize("form", ["adjoint_fp32", "adjoint_mixed"])
def test_config1_lowp_forms(orc, cfg1, form):
    c = cfg1
    s, gp, b2 = build(c["cell"], c["species"], c["positions"], c["settings"], 5.0,
                      c["sigma"], 2, c["sparse_atoms"], c["alpha"])
    o, g, P, Pn, m = oracle_side(orc, c["cell"], c["species"], c["positions"],
                                 c["settings"], 5.0, c["sparse_atoms"])
    V = orc.volume(c["cell"])
    want = orc.predict_mean(o, m, P, Pn, c["sigma"], 2, V, c["alpha"])
    Li_o = orc.cholesky_inverse(orc.envs_envs(1, m, P, Pn, m, P, Pn, c["sigma"], 2),
                                1e-8)
    var_o = orc.cluster_uncertainties(o, m, P, Pn, c["sigma"], 2, Li_o)
    n = s.noa
    efs, (var,) = gp.predict_local_uncertainties(s, form=form)
    e_err = abs(efs[0] - want[0]) / abs(want[0])
    f_err = nrel(efs[1:1 + 3 * n], want[1:1 + 3 * n])
    v_err = np.abs(var - var_o).max() / c["sigma"] ** 2
    print(f"{form}: energy rel {e_err:.3e}, force rel {f_err:.3e}, "
          f"variance {v_err:.3e} sigma^2", flush=True)
    tol = LOWP_TOL[form]
    assert e_err < tol[0] and f_err < tol[1] and v_err < tol[2]


@pytest.mark.parametrize("form", ["adjoint_fp32", "adjoint_mixed"])
def test_config3_lowp_mean(form):
    # 32,000 atoms, 500 sparse: the low-precision forms against the FP64 GPU
    # form (itself pinned to the oracle above); errors in meV/atom, meV/A
    c = synth.config3()
    s, gp, b2 = build(c["cell"], c["species"], c["positions"], c["settings"], 5.0,
                      c["sigma"], c["power"], c["sparse_atoms"], c["alpha"])
    n = s.noa
    ref = gp.predict_mean(s, form="adjoint_fp64").copy()
    efs = gp.predict_mean(s, form=form)
    de = abs(efs[0] - ref[0]) / n * 1e3
    df = (efs[1:1 + 3 * n] - ref[1:1 + 3 * n]) * 1e3
    print(f"{form}: energy {de:.3e} meV/atom, force RMS {np.sqrt(np.mean(df**2)):.3e}"
          f" max {np.abs(df).max():.3e} meV/A (|F| max "
          f"{np.abs(ref[1:1 + 3 * n]).max() * 1e3:.3e})", flush=True)
    assert np.sqrt(np.mean(df ** 2)) < 1.0  # meV/A
    assert nrel(efs[1:1 + 3 * n], ref[1:1 + 3 * n]) < LOWP_TOL[form][1]


def test_config4_variance_mixed():
    # 55,296 atoms, 3 species, 2,000 sparse: split-FP16 kernel matrix against
    # the FP64 path; relative variance error
    c = synth.config4()
    s, gp, b2 = build(c["cell"], c["species"], c["positions"], c["settings"],
                      c["cutoff"], c["sigma"], c["power"], c["sparse_atoms"],
                      c["alpha"])
    efs64, (var64,) = gp.predict_local_uncertainties(s, form="adjoint_fp64")
    efs64 = efs64.copy()
    var64 = var64.copy()
    s2 = c["sigma"] ** 2
    for form in ("adjoint_fp32", "adjoint_mixed"):
        efs, (var,) = gp.predict_local_uncertainties(s, form=form)
        n = s.noa
        dv = np.abs(var - var64)
        rel = dv / np.maximum(np.abs(var64), 1e-12)
        # relative error by variance band: std / sigma in [0.03, 0.1) and
        # [0.1, 1], where uncertainty thresholds are set
        small = (var64 >= 1e-3 * s2) & (var64 < 1e-2 * s2)
        large = var64 >= 1e-2 * s2
        print(f"{form}: variance abs max {dv.max():.3e} (sigma^2 = {s2}), rel "
              f"median {np.median(rel):.3e}; std/sigma in [0.03, 0.1): "
              f"{small.sum()} atoms, rel max {rel[small].max():.3e}; >= 0.1: "
              f"{large.sum()} atoms, rel max {rel[large].max():.3e}; energy "
              f"{abs(efs[0] - efs64[0]) / n * 1e3:.3e} meV/atom", flush=True)
        assert dv.max() < LOWP_TOL[form][2] * s2
        assert rel[small].max() < 2e-2 and rel[large].max() < 2e-3, form
